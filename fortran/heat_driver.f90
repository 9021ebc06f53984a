! heat_driver.f90 -- the heat problem driven through the drop-in modules,
! like problems/heat/main.F90 on one rank:
!     heat_driver SIZE ORDER steps Dt [f0] [out.bin]
! f0: 0 constant, 1 prod sin(pi x_i), 2 the heat problem's 2*bump3d (default).
! Prints "iter L2 norm: value" per step (main.F90:55) and optionally writes
! the final coefficients (x fastest, raw float64).
module heat_rhs
   use Setup
   implicit none
   real(kind=8) :: Dt = 0.d0, t = 0.d0
contains
   ! host statement of the heat integrand (problems/heat/RHS_eq.F90:35-88);
   ! Step runs its registered device form instead of calling it
   subroutine HeatPoint(ads, X, k, e, a, du, n, un, un13, un23, ads_data, J, W, direction, substep, l2norm, ret)
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
      real(kind=8) :: v, dvx, dvy, dvz
      v = ads%NNx(0, a(1), k(1), e(1)) * ads%NNy(0, a(2), k(2), e(2)) * ads%NNz(0, a(3), k(3), e(3))
      dvx = ads%NNx(1, a(1), k(1), e(1)) * ads%NNy(0, a(2), k(2), e(2)) * ads%NNz(0, a(3), k(3), e(3))
      dvy = ads%NNx(0, a(1), k(1), e(1)) * ads%NNy(1, a(2), k(2), e(2)) * ads%NNz(0, a(3), k(3), e(3))
      dvz = ads%NNx(0, a(1), k(1), e(1)) * ads%NNy(0, a(2), k(2), e(2)) * ads%NNz(1, a(3), k(3), e(3))
      ret = J * W * (v * un(1) - Dt * (du(1) * dvx + du(2) * dvy + du(3) * dvz))
      l2norm = J * W * v * un(1) * un(1)
   end subroutine HeatPoint
end module heat_rhs

program heat_driver
   use iso_c_binding
   use Setup
   use ADSS
   use heat_rhs
   implicit none
   type(ADS_setup) :: ads
   type(ADS_compute_data) :: ads_data
   integer(kind=4) :: sz, order, steps, iter, mierr, f0, nargs
   real(kind=8) :: l2norm
   character(len=256) :: arg, outfile

   nargs = command_argument_count()
   if (nargs < 4) then
      write(*, '(a)') 'usage: heat_driver SIZE ORDER steps Dt [f0] [out.bin]'
      stop 5
   end if
   call get_command_argument(1, arg); read(arg, *) sz
   call get_command_argument(2, arg); read(arg, *) order
   call get_command_argument(3, arg); read(arg, *) steps
   call get_command_argument(4, arg); read(arg, *) Dt
   f0 = ADS_F0_BUMP3D
   if (nargs >= 5) then
      call get_command_argument(5, arg); read(arg, *) f0
   end if
   outfile = ''
   if (nargs >= 6) call get_command_argument(6, outfile)

   call initialize((/ sz, sz, sz /), (/ order, order, order /), ads, ads_data, mierr)
   if (mierr /= 0) stop 2
   call RegisterHeatForm(HeatPoint, Dt, int(f0, c_int), mierr)
   t = 0.d0
   do iter = 0, steps
      call Step(iter, HeatPoint, ads, ads_data, l2norm, mierr)
      if (mierr /= 0) stop 3
      write(*, '(i6,a,es24.16)') iter, ' L2 norm:', l2norm
      t = t + Dt
      call PrintSolution(iter, t, ads, ads_data)
   end do
   if (len_trim(outfile) > 0) then
      open(unit=10, file=trim(outfile), access='stream', form='unformatted', status='replace')
      write(10) ads_data%F
      close(10)
   end if
   call Cleanup(ads, ads_data, mierr)
end program heat_driver
