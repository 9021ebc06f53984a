! oil_driver.f90 -- the oil problem driven through the drop-in modules, like
! problems/oil/main.F90 on one rank:
!     oil_driver SIZE ORDER steps Dt seed npumps (x y z)... ndrains (x y z)... [out.bin]
! The pumps and drains are read as InitPumps reads them (input_data.F90:99-136).
! The curves follow InitInputData's random walk with a seeded stream
! (ads_oil_curves; the reference's random_number is unseeded); the constants
! are input_data's (mi = 10, Kq 1..1000, radius 0.15, strengths 1).  Prints
! "iter L2 norm: value" per step, then the drained total (ComputeResults),
! and optionally writes the final coefficients (x fastest, raw float64).
module oil_rhs
   use Setup
   implicit none
   integer(kind=4), parameter :: cN = 30, cL = 16
   real(kind=8), parameter :: mi = 10.d0, Kqmin = 1.d0, Kqmax = 1000.d0
   real(kind=8), parameter :: radius = 0.15d0, pumping_strength = 1.d0, draining_strength = 1.d0
   real(kind=8) :: Dt = 0.d0, t = 0.d0
contains
   ! host statement of the oil integrand's t > 0 branch (problems/oil/RHS_eq.F90:76-87)
   ! without the sources; Step runs the registered device form instead
   subroutine OilPoint(ads, X, k, e, a, du, n, un, un13, un23, ads_data, J, W, direction, substep, l2norm, ret)
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
      ret = J * W * (v * un(1) - Dt * exp(mi * un(1)) * (du(1) * dvx + du(2) * dvy + du(3) * dvz))
      l2norm = J * W * v * un(1) * un(1)
   end subroutine OilPoint
end module oil_rhs

program oil_driver
   use iso_c_binding
   use Setup
   use ADSS
   use ads_capi, only: ads_oil_curves
   use oil_rhs
   implicit none
   type(ADS_setup) :: ads
   type(ADS_compute_data) :: ads_data
   integer(kind=4) :: sz, order, steps, iter, mierr, nargs, np, nd, i, arg_i
   integer(kind=8) :: seed
   real(kind=8) :: l2norm
   real(kind=8), allocatable :: pumps(:, :), drains(:, :)
   real(c_double) :: cx(cN * cL), cy(cN * cL), cz(cN * cL)
   character(len=256) :: arg, outfile

   nargs = command_argument_count()
   if (nargs < 7) then
      write(*, '(a)') 'usage: oil_driver SIZE ORDER steps Dt seed npumps (x y z)... ndrains (x y z)... [out.bin]'
      stop 5
   end if
   call get_command_argument(1, arg); read(arg, *) sz
   call get_command_argument(2, arg); read(arg, *) order
   call get_command_argument(3, arg); read(arg, *) steps
   call get_command_argument(4, arg); read(arg, *) Dt
   call get_command_argument(5, arg); read(arg, *) seed
   arg_i = 6
   call get_command_argument(arg_i, arg); read(arg, *) np
   allocate(pumps(3, np))
   do i = 1, np
      call get_command_argument(arg_i + 1, arg); read(arg, *) pumps(1, i)
      call get_command_argument(arg_i + 2, arg); read(arg, *) pumps(2, i)
      call get_command_argument(arg_i + 3, arg); read(arg, *) pumps(3, i)
      arg_i = arg_i + 3
   end do
   arg_i = arg_i + 1
   call get_command_argument(arg_i, arg); read(arg, *) nd
   allocate(drains(3, nd))
   do i = 1, nd
      call get_command_argument(arg_i + 1, arg); read(arg, *) drains(1, i)
      call get_command_argument(arg_i + 2, arg); read(arg, *) drains(2, i)
      call get_command_argument(arg_i + 3, arg); read(arg, *) drains(3, i)
      arg_i = arg_i + 3
   end do
   outfile = ''
   if (nargs > arg_i) call get_command_argument(arg_i + 1, outfile)

   call initialize((/ sz, sz, sz /), (/ order, order, order /), ads, ads_data, mierr)
   if (mierr /= 0) stop 2
   mierr = ads_oil_curves(int(seed, c_int64_t), cN, cL, cx, cy, cz)
   if (mierr /= 0) stop 2
   call RegisterOilForm(OilPoint, Dt, ads, cx, cy, cz, cN, cL, pumps, drains, mi, Kqmin, Kqmax, radius, &
                        pumping_strength, draining_strength, mierr)
   if (mierr /= 0) stop 2
   t = 0.d0
   do iter = 0, steps
      call Step(iter, OilPoint, ads, ads_data, l2norm, mierr)
      if (mierr /= 0) stop 3
      write(*, '(i6,a,es24.16)') iter, ' L2 norm:', l2norm
      t = t + Dt
   end do
   write(*, '(a,es24.16)') 'drained:', oil_drained
   if (len_trim(outfile) > 0) then
      open(unit=10, file=trim(outfile), access='stream', form='unformatted', status='replace')
      write(10) ads_data%F
      close(10)
   end if
   call Cleanup(ads, ads_data, mierr)
end program oil_driver
