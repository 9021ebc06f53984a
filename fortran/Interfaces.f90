! Interfaces.f90 -- drop-in for module Interfaces (src/Interfaces.F90:52-81):
! the right-hand-side callback interface, argument for argument.
module Interfaces
   interface
      subroutine RHS_fun_int(ads, X, k, e, a, du, n, un, un13, un23, ads_data, J, W, direction, substep, &
                             l2norm, ret)
         use Setup
         implicit none
         type(ADS_setup), intent(in) :: ads
         real(kind=8), intent(in), dimension(3) :: X
         integer(kind=4), intent(in), dimension(3) :: k, e, a
         real(kind=8), intent(in), dimension(3) :: du
         integer(kind=4), intent(in) :: n
         real(kind=8), intent(in), dimension(n) :: un
         real(kind=8), intent(in) :: un13, un23
         type(ADS_compute_data), intent(in) :: ads_data
         real(kind=8), intent(in) :: J, W
         integer(kind=4), intent(in) :: direction, substep
         real(kind=8), intent(out) :: l2norm, ret
      end subroutine
   end interface
end module Interfaces
