! projection_engine.f90 -- drop-in for the entry points of module
! projection_engine (src/projection_engine.F90) on the GPU path.
module projection_engine
   use iso_c_binding
   use Setup, only: ADS_setup, ADS_compute_data
   use Interfaces, only: RHS_fun_int
   use ads_capi
   implicit none
   private
   public :: FormUn, Form3DRHS, ComputeMatrix

contains

   ! FormUn (:306-441): values and gradients at the Gauss points are computed
   ! inside the assembly kernel and never stored, so there is nothing to do.
   subroutine FormUn(subun, ads, ads_data)
      integer(kind=4), intent(in) :: subun
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      if (subun < 1 .or. subun > 3) write(*, '(a)') 'wrong substep'
   end subroutine FormUn

   ! Form3DRHS (:154-284) for the heat integrand at t > 0 with time step Dt:
   ! ads_data%F receives the right-hand side for the current coefficients.
   subroutine Form3DRHS(ads, ads_data, direction, substep, Dt, l2norm)
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      integer(kind=4), intent(in) :: direction, substep
      real(kind=8), intent(in) :: Dt
      real(kind=8), intent(out) :: l2norm
      integer(c_int) :: rc
      interface
         integer(c_int) function ads_form_rhs(ctx, dt, F, l2) bind(C, name="ads_form_rhs")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
            real(c_double), intent(out) :: F(*)
            real(c_double), intent(out) :: l2
         end function
      end interface
      rc = ads_form_rhs(ads%handle, Dt, ads_data%F, l2norm)
      if (rc /= ADS_OK) write(*, '(a,a)') 'Form3DRHS: ', ads_error_text(ads%handle)
   end subroutine Form3DRHS

   ! ComputeMatrix / MKBBT (:531-544, :63-132) as a dense (n+1)x(n+1) matrix
   ! of direction dir (1..3); the reference's sparse linked list is not kept.
   subroutine ComputeMatrix(ads, dir, mix, A)
      type(ADS_setup), intent(in) :: ads
      integer(kind=4), intent(in) :: dir
      real(kind=8), intent(in) :: mix(4)
      real(kind=8), intent(out) :: A(:, :)
      real(c_double), allocatable :: rowmajor(:)
      integer(c_int) :: rc
      integer :: nn, i
      nn = ads%n(dir) + 1
      allocate(rowmajor(nn * nn))
      rc = ads_get_matrix(ads%handle, int(dir - 1, c_int), mix, rowmajor)
      do i = 1, nn
         A(i, 1:nn) = rowmajor((i - 1) * nn + 1:i * nn)
      end do
   end subroutine ComputeMatrix

end module projection_engine
