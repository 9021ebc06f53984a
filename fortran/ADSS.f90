! ADSS.f90 -- drop-in for module ADSS (src/ADS.F90): initialize, Step,
! MultiStep, Sub_Step, SolveOneDirection, Cleanup and PrintSolution keep
! their names and arguments; the work runs on the GPU through the C ABI
! (include/ads_mi355x.h).
!
! The per-point callback RHS_fun cannot run on the device.  A problem binds
! its callback to a device form once, after initialize:
!     call RegisterHeatForm(ComputePointForRHS, Dt, ADS_F0_BUMP3D, mierr, t=t)
!     call RegisterProjectionForm(ComputePointForRHS, ADS_F0_CONST, mierr)
!     call RegisterOilForm(ComputePointForRHS, Dt, ads, cx, cy, cz, cN, cL, &
!                          pumps, drains, mi, Kqmin, Kqmax, radius, ps, ds, mierr)
! and the entry points look the procedure up by address (c_funloc).  An
! unregistered procedure makes them return mierr = ADS_ERR_UNSUPPORTED.
module ADSS
   use iso_c_binding
   use Setup, only: ADS_setup, ADS_compute_data
   use Interfaces, only: RHS_fun_int
   use ads_capi
   use ads_forms
   implicit none
   private

   public :: initialize, Step, MultiStep, Sub_Step, SolveOneDirection, Cleanup, PrintSolution
   public :: RegisterHeatForm, RegisterProjectionForm, RegisterOilForm, SyncSolution, SetDiffusivity
   public :: ads_sync_every_step, oil_drained
   public :: ADS_F0_CONST, ADS_F0_SINPROD, ADS_F0_BUMP3D, ADS_PREC_FP64, ADS_PREC_FP32_IR, ADS_PREC_FP32_IR_KRON, &
             ADS_PREC_FP64_KRON, ADS_PREC_FP32_IR_SEG

   ! copy the coefficients back into ads_data%F after every Step (the
   ! reference's data contract); set .false. and call SyncSolution to read
   ! them only when needed
   logical, save :: ads_sync_every_step = .true.

   ! the oil form's running total of J W v Dt vdrain over its steps (the
   ! problem's module variable `drained`, problems/oil/RHS_eq.F90:86)
   real(kind=8), save :: oil_drained = 0.d0

contains

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

   ! Binds a heat integrand (problems/heat/RHS_eq.F90:35-88) to the device
   ! form: its t == 0 branch projects f0, its t > 0 branch takes one explicit
   ! step of length Dt.  With t (the problem's own time variable, input_data's
   ! t) the branch follows t as the integrand's test does; without it Step
   ! takes the t == 0 branch at iter == 0.
   subroutine RegisterHeatForm(RHS_fun, Dt, f0, mierr, prec, refine, t)
      procedure(RHS_fun_int) :: RHS_fun
      real(kind=8), intent(in) :: Dt
      integer(c_int), intent(in) :: f0
      integer(kind=4), intent(out) :: mierr
      integer(c_int), intent(in), optional :: prec, refine
      real(kind=8), intent(in), target, optional :: t
      integer(c_int) :: pr, rf
      type(c_ptr) :: tp
      pr = ADS_PREC_FP64
      rf = 2
      tp = c_null_ptr
      if (present(prec)) pr = prec
      if (present(refine)) rf = refine
      if (present(t)) tp = c_loc(t)
      call add_form(RHS_fun, FORM_HEAT, f0, real(Dt, c_double), pr, rf, tp, mierr)
   end subroutine

   ! Binds a projection integrand (problems/L2Projection/RHS_eq.F90:34-71).
   subroutine RegisterProjectionForm(RHS_fun, f0, mierr)
      procedure(RHS_fun_int) :: RHS_fun
      integer(c_int), intent(in) :: f0
      integer(kind=4), intent(out) :: mierr
      call add_form(RHS_fun, FORM_PROJECTION, f0, 0.d0, ADS_PREC_FP64, 1_c_int, c_null_ptr, mierr)
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
      call add_form(RHS_fun, FORM_OIL, ADS_F0_CONST, real(Dt, c_double), ADS_PREC_FP64, 1_c_int, c_null_ptr, mierr)
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

   ! ADSS::Step (src/ADS.F90:437-461): FormUn(1) and one Sub_Step with mix
   ! (1, 0, 0, 0), i.e. the right-hand side solved with the mass matrix in
   ! x, y and z, as one C-ABI call (the heat form's precision form applies).
   subroutine Step(iter, RHS_fun, ads, ads_data, l2norm, mierr)
      integer(kind=4), intent(in) :: iter
      procedure(RHS_fun_int) :: RHS_fun
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      real(kind=8), intent(out) :: l2norm
      integer(kind=4), intent(out) :: mierr
      type(form_entry) :: f
      real(c_double) :: l2, dr
      integer(c_int) :: rc
      l2norm = 0.d0
      if (.not. find_form(RHS_fun, 'Step', f)) then
         mierr = ADS_ERR_UNSUPPORTED
         return
      end if
      if (f%kind == FORM_OIL) then
         if (iter == 0) then
            rc = ads_oil_project(ads%handle, l2)
            ads_data%t = 0.d0
         else
            rc = ads_oil_step(ads%handle, f%dt, l2, dr)
            if (rc == ADS_OK) oil_drained = oil_drained + dr
            ads_data%t = ads_data%t + f%dt
         end if
      else if (f%kind == FORM_PROJECTION .or. form_at_t0(f, iter, ads_data%t)) then
         rc = ads_project(ads%handle, f%f0, l2)
         ads_data%t = 0.d0
      else
         rc = ads_step(ads%handle, f%dt, f%prec, f%refine, l2)
         ads_data%t = ads_data%t + f%dt
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
   ! initial state, gradients of the latest coefficients).  One C-ABI call:
   ! the same as FormUn(s) and Sub_Step(..., mix(:, s), s, s, ...) for s = 1..3.
   subroutine MultiStep(iter, mix, RHS_fun, ads, ads_data, l2norm, mierr)
      integer(kind=4), intent(in) :: iter
      real(kind=8), intent(in) :: mix(4, 3)
      procedure(RHS_fun_int) :: RHS_fun
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      real(kind=8), intent(out) :: l2norm
      integer(kind=4), intent(out) :: mierr
      type(form_entry) :: f
      real(c_double) :: l2
      integer(c_int) :: rc
      l2norm = 0.d0
      if (.not. find_form(RHS_fun, 'MultiStep', f)) then
         mierr = ADS_ERR_UNSUPPORTED
         return
      end if
      if (f%kind /= FORM_HEAT) then
         write(*, '(a)') 'MultiStep: only the heat form has substeps'
         mierr = ADS_ERR_UNSUPPORTED
         return
      end if
      rc = ads_multistep(ads%handle, mix, f%dt, l2)
      call report(ads%handle, rc, 'MultiStep')
      mierr = rc
      if (rc /= ADS_OK) return
      ads_data%t = ads_data%t + f%dt
      l2norm = l2
      if (ads_sync_every_step) call SyncSolution(ads, ads_data, mierr)
   end subroutine MultiStep

   ! ADSS::Sub_Step (src/ADS.F90:471-794): Form3DRHS with the device form
   ! bound to RHS_fun (values of U_n kept by the last FormUn(1), gradients of
   ! the current coefficients), then the solves with mix(1) M + mix(2) K +
   ! mix(3) B + mix(4) B^T in x, y and z (ComputeMatrix + SolveOneDirection
   ! per direction; each distinct mix is factored once and kept), the result
   ! the new coefficients.  iter = 0 (or the heat form's t == 0 branch)
   ! projects the initial state instead.
   subroutine Sub_Step(ads, iter, mix, direction, substep, RHS_fun, ads_data, l2norm, mierr)
      type(ADS_setup), intent(in) :: ads
      integer(kind=4), intent(in) :: iter
      real(kind=8), intent(in) :: mix(4)
      integer(kind=4), intent(in) :: direction, substep
      procedure(RHS_fun_int) :: RHS_fun
      type(ADS_compute_data), intent(inout) :: ads_data
      real(kind=8), intent(out) :: l2norm
      integer(kind=4), intent(out) :: mierr
      type(form_entry) :: f
      real(c_double) :: l2, dr
      integer(c_int) :: rc, form
      l2norm = 0.d0
      if (.not. find_form(RHS_fun, 'Sub_Step', f)) then
         mierr = ADS_ERR_UNSUPPORTED
         return
      end if
      select case (f%kind)
      case (FORM_OIL)
         form = ADS_FORM_OIL
         if (iter == 0) form = ADS_FORM_OIL_PROJECT
      case (FORM_PROJECTION)
         form = ADS_FORM_PROJECT
      case default
         form = ADS_FORM_HEAT
         if (form_at_t0(f, iter, ads_data%t)) form = ADS_FORM_PROJECT
      end select
      rc = ads_sub_step(ads%handle, form, f%f0, mix, int(direction, c_int), int(substep, c_int), f%dt, l2)
      call report(ads%handle, rc, 'Sub_Step')
      mierr = rc
      if (rc /= ADS_OK) return
      if (form == ADS_FORM_OIL) then
         rc = ads_last_drained(ads%handle, dr)
         if (rc == ADS_OK) oil_drained = oil_drained + dr
      end if
      l2norm = l2
      if (ads_sync_every_step) call SyncSolution(ads, ads_data, mierr)
   end subroutine Sub_Step

   ! ADSS::SolveOneDirection (src/ADS.F90:294-379): RHS(1:n+1, 1:eqnum), each
   ! column one right-hand side, overwritten by the solution of sprsmtrx's
   ! system (ComputeMatrix), on the GPU.  The matrix was factored (banded LU
   ! with partial pivoting) when ComputeMatrix built it; the reference runs
   ! MUMPS' analysis, factorisation and one solve per column here.
   subroutine SolveOneDirection(RHS, eqnum, n, p, sprsmtrx)
      use sparse, only: sparse_matrix, matrix_error_text
      real(kind=8) :: RHS(:, :)
      integer(kind=4) :: eqnum
      integer :: n, p
      type(sparse_matrix), pointer, intent(in) :: sprsmtrx
      integer(c_int) :: rc
      if (.not. associated(sprsmtrx)) then
         write(*, '(a)') 'SolveOneDirection: no matrix (ComputeMatrix failed)'
         return
      end if
      if (sprsmtrx%n /= n .or. sprsmtrx%p /= p .or. size(RHS, 1) /= n + 1 .or. size(RHS, 2) < eqnum) then
         write(*, '(a)') 'SolveOneDirection: RHS must be (n+1) x eqnum for the matrix of ComputeMatrix(..., p, n, ...)'
         return
      end if
      if (eqnum < 1) return
      rc = ads_matrix_solve(sprsmtrx%handle, RHS, int(eqnum, c_int))
      if (rc /= ADS_OK) write(*, '(a,i0,a,a)') 'SolveOneDirection: ADS error ', rc, ': ', matrix_error_text(sprsmtrx)
   end subroutine SolveOneDirection

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
