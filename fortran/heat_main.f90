! heat_main.f90 -- the heat problem's main program (problems/heat/main.F90)
! on the drop-in modules, for one process and one GPU: it uses the problem's
! own input_data and RHS_eq modules unchanged (compiled from the problem's
! sources beside this file), registers the integrand with its device form,
! and runs Step / PrintSolution over the steps as the reference's main does.
! The MPI set-up and the l2norm reduction across ranks have no counterpart
! on one process.
!
!     heat_main SIZE ORDER steps Dt procx procy procz    (the reference's arguments)
!
! With a non-empty ADS_NO_VTK in the environment no step<iter>.vti is written.
program heat_main
   use Setup, only: ADS_setup, ADS_compute_data
   use RHS_eq
   use ADSS
   use input_data
   implicit none
   integer :: iter
   integer(kind=4) :: ierr
   integer(kind=4), dimension(3) :: p1, p2
   type(ADS_setup) :: ads
   type(ADS_compute_data) :: ads_data
   real(kind=8) :: l2norm
   character(len=8) :: novtk
   integer :: nvlen

   t = 0
   call InitializeParameters
   p1 = (/ SIZE, SIZE, SIZE /)
   p2 = (/ ORDER, ORDER, ORDER /)
   call initialize(p1, p2, ads, ads_data, ierr)
   if (ierr /= 0) stop 1
   ! the heat integrand: 2 bump3d(0.05, 0.4) at t = 0 (input_data's
   ! initial_state), kappa = 1 and f = 0 after; its branch follows t
   call RegisterHeatForm(ComputePointForRHS, Dt, ADS_F0_BUMP3D, ierr, t=t)
   if (ierr /= 0) stop 1
   call get_environment_variable('ADS_NO_VTK', novtk, nvlen)

   do iter = 0, steps
      l2norm = 0.d0
      call Step(iter, ComputePointForRHS, ads, ads_data, l2norm, ierr)
      if (ierr /= 0) stop 1
      write(*, *) iter, 'L2 norm:', l2norm
      t = t + Dt
      if (nvlen == 0) call PrintSolution(iter, t, ads, ads_data)
   end do

   call Cleanup(ads, ads_data, ierr)
end program heat_main
