! ADSS.f90 -- drop-in for module ADSS (src/ADS.F90): initialize, Step,
! MultiStep, Cleanup and PrintSolution keep their names and arguments; the
! work runs on the GPU through the C ABI (include/ads_mi355x.h).
!
! The per-point callback RHS_fun cannot run on the device.  A problem binds
! its callback to a device form once, after initialize:
!     call RegisterHeatForm(ComputePointForRHS, Dt, ADS_F0_BUMP3D, mierr)
!     call RegisterProjectionForm(ComputePointForRHS, ADS_F0_CONST, mierr)
!     call RegisterOilForm(ComputePointForRHS, Dt, ads, cx, cy, cz, cN, cL, &
!                          pumps, drains, mi, Kqmin, Kqmax, radius, ps, ds, mierr)
! and Step looks the procedure up by address (c_funloc).  An unregistered
! procedure makes Step return mierr = ADS_ERR_UNSUPPORTED.
module ADSS
   use iso_c_binding
   use Setup, only: ADS_setup, ADS_compute_data
   use Interfaces, only: RHS_fun_int
   use ads_capi
   implicit none
   private

   public :: initialize, Step, MultiStep, Cleanup, PrintSolution
   public :: RegisterHeatForm, RegisterProjectionForm, RegisterOilForm, SyncSolution, SetDiffusivity
   public :: ads_sync_every_step, oil_drained
   public :: ADS_F0_CONST, ADS_F0_SINPROD, ADS_F0_BUMP3D, ADS_PREC_FP64, ADS_PREC_FP32_IR, ADS_PREC_FP32_IR_KRON, &
             ADS_PREC_FP64_KRON, ADS_PREC_FP32_IR_SEG

   integer, parameter :: FORM_HEAT = 1, FORM_PROJECTION = 2, FORM_OIL = 3, MAXFORMS = 16

   type form_entry
      type(c_funptr) :: fptr = c_null_funptr
      integer :: kind = 0
      integer(c_int) :: f0 = ADS_F0_CONST
      real(c_double) :: dt = 0.d0
      integer(c_int) :: prec = ADS_PREC_FP64
      integer(c_int) :: refine = 2
   end type form_entry

   type(form_entry), save :: forms(MAXFORMS)
   integer, save :: nforms = 0

   ! copy the coefficients back into ads_data%F after every Step (the
   ! reference's data contract); set .false. and call SyncSolution to read
   ! them only when needed
   logical, save :: ads_sync_every_step = .true.

   ! the oil form's running total of J W v Dt vdrain over its steps (the
   ! problem's module variable `drained`, problems/oil/RHS_eq.F90:86)
   real(kind=8), save :: oil_drained = 0.d0

contains

   subroutine report(ctx, code, what)
      type(c_ptr), intent(in) :: ctx
      integer(c_int), intent(in) :: code
      character(len=*), intent(in) :: what
      if (code /= ADS_OK) write(*, '(a,a,i0,a,a)') what, ': ADS error ', code, ': ', ads_error_text(ctx)
   end subroutine

   ! ADSS::initialize (src/ADS.F90:11-107) on one rank
   subroutine initialize(n, p, ads, ads_data, mierr)
      integer(kind=4), intent(in), dimension(3) :: n, p
      type(ADS_setup), intent(out) :: ads
      type(ADS_compute_data), intent(out) :: ads_data
      integer(kind=4), intent(out) :: mierr
      integer(c_int) :: dofs(3), ne(3), pp(3), rc
      integer :: d, q

      rc = ads_init(int(n, c_int), int(p, c_int), ads%handle)
      if (rc /= ADS_OK) then
         call report(c_null_ptr, rc, 'initialize')
         mierr = rc
         return
      end if
      rc = ads_get_sizes(ads%handle, dofs, ne, pp)
      ads%n = n
      ads%p = p
      ads%nelem = ne
      ads%m = n + p + 1
      ads%ng = p + 1
      ! one rank: the whole domain (parallelism.F90:146-171 with rank 0 of 1)
      ads%nrcpp = n + 1
      ads%ibeg = 1
      ads%iend = n + 1
      ads%s = n + 1
      ads%mine = 1
      ads%maxe = ne
      ads%lnelem = ne
      ads%ibegsx = -1; ads%iendsx = -1
      ads%ibegsy = -1; ads%iendsy = -1
      ads%ibegsz = -1; ads%iendsz = -1
      ads%ibegsx(2) = 1; ads%iendsx(2) = n(1) + 1
      ads%ibegsy(2) = 1; ads%iendsy(2) = n(2) + 1
      ads%ibegsz(2) = 1; ads%iendsz(2) = n(3) + 1
      do d = 1, 3
         q = p(d) + 1
         select case (d)
         case (1)
            allocate(ads%Ux(n(1) + p(1) + 2), ads%Ox(ne(1)), ads%Jx(ne(1)), ads%Wx(q), ads%Xx(q, ne(1)), &
                     ads%NNx(0:1, 0:p(1), q, ne(1)))
            rc = ads_get_knots(ads%handle, 0, ads%Ux)
            rc = ads_get_basis(ads%handle, 0, ads%Ox, ads%Jx, ads%Wx, ads%Xx, ads%NNx)
         case (2)
            allocate(ads%Uy(n(2) + p(2) + 2), ads%Oy(ne(2)), ads%Jy(ne(2)), ads%Wy(q), ads%Xy(q, ne(2)), &
                     ads%NNy(0:1, 0:p(2), q, ne(2)))
            rc = ads_get_knots(ads%handle, 1, ads%Uy)
            rc = ads_get_basis(ads%handle, 1, ads%Oy, ads%Jy, ads%Wy, ads%Xy, ads%NNy)
         case (3)
            allocate(ads%Uz(n(3) + p(3) + 2), ads%Oz(ne(3)), ads%Jz(ne(3)), ads%Wz(q), ads%Xz(q, ne(3)), &
                     ads%NNz(0:1, 0:p(3), q, ne(3)))
            rc = ads_get_knots(ads%handle, 2, ads%Uz)
            rc = ads_get_basis(ads%handle, 2, ads%Oz, ads%Jz, ads%Wz, ads%Xz, ads%NNz)
         end select
      end do
      allocate(ads_data%F(ads%s(1), ads%s(2) * ads%s(3)))
      ads_data%F = 0.d0
      ads_data%t = 0.d0
      mierr = 0
   end subroutine initialize

   subroutine add_form(RHS_fun, kind, f0, dt, prec, refine, mierr)
      procedure(RHS_fun_int) :: RHS_fun
      integer, intent(in) :: kind
      integer(c_int), intent(in) :: f0, prec, refine
      real(c_double), intent(in) :: dt
      integer(kind=4), intent(out) :: mierr
      type(c_funptr) :: fp
      integer :: i
      fp = c_funloc(RHS_fun)
      do i = 1, nforms
         if (c_associated(forms(i)%fptr, fp)) exit
      end do
      if (i > nforms) then
         if (nforms == MAXFORMS) then
            mierr = ADS_ERR_INVALID
            return
         end if
         nforms = nforms + 1
         i = nforms
      end if
      forms(i) = form_entry(fp, kind, f0, dt, prec, refine)
      mierr = 0
   end subroutine

   ! Binds a heat integrand (problems/heat/RHS_eq.F90:35-88) to the device
   ! form: Step(0, ...) projects f0 (t = 0 branch), Step(iter > 0, ...) takes
   ! one explicit step of length Dt (t > 0 branch).
   subroutine RegisterHeatForm(RHS_fun, Dt, f0, mierr, prec, refine)
      procedure(RHS_fun_int) :: RHS_fun
      real(kind=8), intent(in) :: Dt
      integer(c_int), intent(in) :: f0
      integer(kind=4), intent(out) :: mierr
      integer(c_int), intent(in), optional :: prec, refine
      integer(c_int) :: pr, rf
      pr = ADS_PREC_FP64
      rf = 2
      if (present(prec)) pr = prec
      if (present(refine)) rf = refine
      call add_form(RHS_fun, FORM_HEAT, f0, real(Dt, c_double), pr, rf, mierr)
   end subroutine

   ! Binds a projection integrand (problems/L2Projection/RHS_eq.F90:34-71).
   subroutine RegisterProjectionForm(RHS_fun, f0, mierr)
      procedure(RHS_fun_int) :: RHS_fun
      integer(c_int), intent(in) :: f0
      integer(kind=4), intent(out) :: mierr
      call add_form(RHS_fun, FORM_PROJECTION, f0, 0.d0, ADS_PREC_FP64, 1_c_int, mierr)
   end subroutine

   ! Binds the oil integrand (problems/oil/RHS_eq.F90:35-94) to the device
   ! form and does InitInputData's and PrecomputeKq's work on the device:
   ! curves cx, cy, cz (cN curves of cL points, curve i at (i-1)*cL+1..i*cL),
   ! pumps(3, np) and drains(3, nd) as InitPumps reads them, and the
   ! input_data constants.  Step(0, ...) projects initial_state, Step(iter >
   ! 0, ...) takes one explicit step of length Dt and adds its drained to
   ! oil_drained.
   subroutine RegisterOilForm(RHS_fun, Dt, ads, cx, cy, cz, cN, cL, pumps, drains, mi, Kqmin, Kqmax, radius, &
                              pumping_strength, draining_strength, mierr)
      procedure(RHS_fun_int) :: RHS_fun
      real(kind=8), intent(in) :: Dt
      type(ADS_setup), intent(in) :: ads
      real(kind=8), intent(in) :: cx(:), cy(:), cz(:), pumps(:, :), drains(:, :)
      integer(kind=4), intent(in) :: cN, cL
      real(kind=8), intent(in) :: mi, Kqmin, Kqmax, radius, pumping_strength, draining_strength
      integer(kind=4), intent(out) :: mierr
      real(c_double) :: prm(6)
      real(c_double) :: none(3)
      prm = (/ mi, Kqmin, Kqmax, radius, pumping_strength, draining_strength /)
      none = 0.d0
      if (size(pumps, 1) /= 3 .or. size(drains, 1) /= 3 .or. size(cx) < cN * cL .or. size(cy) < cN * cL &
          .or. size(cz) < cN * cL) then
         mierr = ADS_ERR_INVALID
         write(*, '(a)') 'RegisterOilForm: curves need cN*cL points, pumps and drains 3 rows'
         return
      end if
      if (size(pumps, 2) > 0 .and. size(drains, 2) > 0) then
         mierr = ads_oil_setup(ads%handle, cx, cy, cz, int(cN, c_int), int(cL, c_int), pumps, &
                               int(size(pumps, 2), c_int), drains, int(size(drains, 2), c_int), prm)
      else if (size(pumps, 2) > 0) then
         mierr = ads_oil_setup(ads%handle, cx, cy, cz, int(cN, c_int), int(cL, c_int), pumps, &
                               int(size(pumps, 2), c_int), none, 0_c_int, prm)
      else if (size(drains, 2) > 0) then
         mierr = ads_oil_setup(ads%handle, cx, cy, cz, int(cN, c_int), int(cL, c_int), none, 0_c_int, &
                               drains, int(size(drains, 2), c_int), prm)
      else
         mierr = ads_oil_setup(ads%handle, cx, cy, cz, int(cN, c_int), int(cL, c_int), none, 0_c_int, &
                               none, 0_c_int, prm)
      end if
      call report(ads%handle, mierr, 'RegisterOilForm')
      if (mierr /= ADS_OK) return
      call add_form(RHS_fun, FORM_OIL, ADS_F0_CONST, real(Dt, c_double), ADS_PREC_FP64, 1_c_int, mierr)
   end subroutine

   ! per-element diffusivity (nelem(1)*nelem(2)*nelem(3), ex fastest)
   subroutine SetDiffusivity(ads, kappa, mierr)
      type(ADS_setup), intent(in) :: ads
      real(kind=8), intent(in), target :: kappa(:)
      integer(kind=4), intent(out) :: mierr
      mierr = ads_set_kappa(ads%handle, c_loc(kappa))
      call report(ads%handle, mierr, 'SetDiffusivity')
   end subroutine

   subroutine SyncSolution(ads, ads_data, mierr)
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      integer(kind=4), intent(out) :: mierr
      mierr = ads_get_coeffs(ads%handle, ads_data%F)
      call report(ads%handle, mierr, 'SyncSolution')
   end subroutine

   ! ADSS::Step (src/ADS.F90:437-461)
   subroutine Step(iter, RHS_fun, ads, ads_data, l2norm, mierr)
      integer(kind=4), intent(in) :: iter
      procedure(RHS_fun_int) :: RHS_fun
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      real(kind=8), intent(out) :: l2norm
      integer(kind=4), intent(out) :: mierr
      type(c_funptr) :: fp
      integer :: i
      real(c_double) :: l2, dr
      integer(c_int) :: rc
      fp = c_funloc(RHS_fun)
      l2norm = 0.d0
      do i = 1, nforms
         if (c_associated(forms(i)%fptr, fp)) exit
      end do
      if (i > nforms) then
         write(*, '(a)') 'Step: RHS_fun is not bound to a device form (RegisterHeatForm/RegisterOilForm/...)'
         mierr = ADS_ERR_UNSUPPORTED
         return
      end if
      if (forms(i)%kind == FORM_OIL) then
         if (iter == 0) then
            rc = ads_oil_project(ads%handle, l2)
            ads_data%t = 0.d0
         else
            rc = ads_oil_step(ads%handle, forms(i)%dt, l2, dr)
            if (rc == ADS_OK) oil_drained = oil_drained + dr
            ads_data%t = ads_data%t + forms(i)%dt
         end if
      else if (forms(i)%kind == FORM_PROJECTION .or. iter == 0) then
         rc = ads_project(ads%handle, forms(i)%f0, l2)
         ads_data%t = 0.d0
      else
         rc = ads_step(ads%handle, forms(i)%dt, forms(i)%prec, forms(i)%refine, l2)
         ads_data%t = ads_data%t + forms(i)%dt
      end if
      call report(ads%handle, rc, 'Step')
      mierr = rc
      if (rc /= ADS_OK) return
      l2norm = l2
      if (ads_sync_every_step) call SyncSolution(ads, ads_data, mierr)
   end subroutine Step

   ! ADSS::MultiStep (src/ADS.F90:391-429): three substeps, substep s solving
   ! with mix(1,s)M + mix(2,s)K + mix(3,s)B + mix(4,s)B^T in every direction;
   ! the registered heat form supplies the integrand (values of the step's
   ! initial state, gradients of the latest coefficients).
   subroutine MultiStep(iter, mix, RHS_fun, ads, ads_data, l2norm, mierr)
      integer(kind=4), intent(in) :: iter
      real(kind=8), intent(in) :: mix(4, 3)
      procedure(RHS_fun_int) :: RHS_fun
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      real(kind=8), intent(out) :: l2norm
      integer(kind=4), intent(out) :: mierr
      type(c_funptr) :: fp
      integer :: i
      real(c_double) :: l2
      integer(c_int) :: rc
      fp = c_funloc(RHS_fun)
      l2norm = 0.d0
      do i = 1, nforms
         if (c_associated(forms(i)%fptr, fp)) exit
      end do
      if (i > nforms) then
         write(*, '(a)') 'MultiStep: RHS_fun is not bound to a device form (RegisterHeatForm)'
         mierr = ADS_ERR_UNSUPPORTED
         return
      end if
      if (forms(i)%kind /= FORM_HEAT) then
         write(*, '(a)') 'MultiStep: only the heat form has substeps'
         mierr = ADS_ERR_UNSUPPORTED
         return
      end if
      rc = ads_multistep(ads%handle, mix, forms(i)%dt, l2)
      call report(ads%handle, rc, 'MultiStep')
      mierr = rc
      if (rc /= ADS_OK) return
      ads_data%t = ads_data%t + forms(i)%dt
      l2norm = l2
      if (ads_sync_every_step) call SyncSolution(ads, ads_data, mierr)
   end subroutine MultiStep

   ! ADSS::PrintSolution (src/ADS.F90:955-989): refreshes ads_data%F and
   ! writes step<iter>.vti, the spline sampled on the 31^3 grid of [0,1]^3
   ! (PlotParams(0,1,0,1,0,1,31,31,31), EvalSpline on the GPU, VtkOutput).
   subroutine PrintSolution(iter, t, ads, ads_data)
      use plot, only: PlotParams
      use vtk, only: VtkOutput
      integer(kind=4), intent(in) :: iter
      real(kind=8), intent(in) :: t
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      integer(kind=4) :: mierr
      type(PlotParams) :: params
      real(kind=8), allocatable :: vals(:, :, :)
      character(len=20) :: filename
      integer(c_int) :: rc
      if (.not. ads_sync_every_step) call SyncSolution(ads, ads_data, mierr)
      params = PlotParams(0.d0, 1.d0, 0.d0, 1.d0, 0.d0, 1.d0, 31, 31, 31)
      allocate(vals(params%resx, params%resy, params%resz))
      rc = ads_eval_grid(ads%handle, [params%resx, params%resy, params%resz], &
                         [params%startx, params%starty, params%startz], [params%endx, params%endy, params%endz], vals)
      if (rc /= 0) then
         write(*, '(a)') 'PrintSolution: '//ads_error_text(ads%handle)
         return
      end if
      write(filename, '(i10)') iter
      filename = 'step'//adjustl(filename)
      call VtkOutput(trim(filename), vals, params)
   end subroutine PrintSolution

   ! ADSS::Cleanup (src/ADS.F90:802-863)
   subroutine Cleanup(ads, ads_data, mierr)
      type(ADS_setup), intent(inout) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      integer(kind=4), intent(out) :: mierr
      mierr = 0
      if (c_associated(ads%handle)) mierr = ads_destroy(ads%handle)
      ads%handle = c_null_ptr
      if (allocated(ads%Ux)) deallocate(ads%Ux, ads%Ox, ads%Jx, ads%Wx, ads%Xx, ads%NNx)
      if (allocated(ads%Uy)) deallocate(ads%Uy, ads%Oy, ads%Jy, ads%Wy, ads%Xy, ads%NNy)
      if (allocated(ads%Uz)) deallocate(ads%Uz, ads%Oz, ads%Jz, ads%Wz, ads%Xz, ads%NNz)
      if (allocated(ads_data%F)) deallocate(ads_data%F)
   end subroutine Cleanup

end module ADSS
