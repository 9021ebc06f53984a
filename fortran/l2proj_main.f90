! l2proj_main.f90 -- the L2-projection problem's main program
! (problems/L2Projection/main.F90) on the drop-in modules, for one process
! and one GPU: the problem's input_data and RHS_eq modules are used
! unchanged; one Step at iter = 0 projects f = 1, and the check is the
! reference's own: l2norm = 1 and every coefficient 1 within epsilon.
!
!     l2proj_main isizex isizey isizez order procx procy procz
program l2proj_main
   use Setup, only: ADS_setup, ADS_compute_data
   use RHS_eq
   use ADSS
   use input_data
   implicit none
   integer(kind=4) :: iter, i, j, ierr
   integer(kind=4), dimension(3) :: p1, p2
   logical :: ok
   type(ADS_setup) :: ads
   type(ADS_compute_data) :: ads_data
   real(kind=8) :: epsilon = 1.e-10
   real(kind=8) :: l2norm

   call InitializeParameters
   p1 = (/ isizex, isizey, isizez /)
   p2 = (/ order, order, order /)
   call initialize(p1, p2, ads, ads_data, ierr)
   if (ierr /= 0) stop 1
   call RegisterProjectionForm(ComputePointForRHS, ADS_F0_CONST, ierr)
   if (ierr /= 0) stop 1

   iter = 0
   l2norm = 0.d0
   call Step(iter, ComputePointForRHS, ads, ads_data, l2norm, ierr)
   if (ierr /= 0) stop 1
   write(*, *) 'L2 norm:', l2norm
   write(*, *) (abs(l2norm - 1.d0) < epsilon)
   ok = .true.
   do i = 1, ads%s(1)
      do j = 1, ads%s(2) * ads%s(3)
         if (abs(ads_data%F(i, j) - 1.d0) > epsilon) ok = .false.
      end do
   end do
   if (.not. ok) write(*, *) 'not OK'
   call Cleanup(ads, ads_data, ierr)
end program l2proj_main
