! ads_capi.f90 -- ISO_C_BINDING interfaces to the C ABI in include/ads_mi355x.h.
module ads_capi
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: ADS_OK = 0, ADS_ERR_INVALID = 1, ADS_ERR_HIP = 2, ADS_ERR_NOMEM = 3, &
      ADS_ERR_SINGULAR = 4, ADS_ERR_UNSUPPORTED = 5, ADS_ERR_NUMERIC = 6
   integer(c_int), parameter :: ADS_F0_CONST = 0, ADS_F0_SINPROD = 1, ADS_F0_BUMP3D = 2
   integer(c_int), parameter :: ADS_PREC_FP64 = 0, ADS_PREC_FP32_IR = 1, ADS_PREC_FP32_IR_KRON = 2, &
                                 ADS_PREC_FP64_KRON = 3, ADS_PREC_FP32_IR_SEG = 4
   integer(c_int), parameter :: ADS_FORM_HEAT = 0, ADS_FORM_PROJECT = 1, ADS_FORM_OIL = 2, ADS_FORM_OIL_PROJECT = 3

   interface
      integer(c_int) function ads_init(n, p, ctx) bind(C, name="ads_init")
         import :: c_int, c_ptr
         integer(c_int), intent(in) :: n(3), p(3)
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int) function ads_destroy(ctx) bind(C, name="ads_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function ads_last_error(ctx) bind(C, name="ads_last_error")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function ads_get_sizes(ctx, dofs, nelem, p) bind(C, name="ads_get_sizes")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: dofs(3), nelem(3), p(3)
      end function
      integer(c_int) function ads_get_knots(ctx, dim, U) bind(C, name="ads_get_knots")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: dim
         real(c_double), intent(out) :: U(*)
      end function
      integer(c_int) function ads_get_basis(ctx, dim, O, J, W, X, NN) bind(C, name="ads_get_basis")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: dim
         integer(c_int), intent(out) :: O(*)
         real(c_double), intent(out) :: J(*), W(*), X(*), NN(*)
      end function
      integer(c_int) function ads_get_matrix(ctx, dim, mix, dense) bind(C, name="ads_get_matrix")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: dim
         real(c_double), intent(in) :: mix(4)
         real(c_double), intent(out) :: dense(*)
      end function
      integer(c_int) function ads_set_coeffs(ctx, host) bind(C, name="ads_set_coeffs")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: host(*)
      end function
      integer(c_int) function ads_get_coeffs(ctx, host) bind(C, name="ads_get_coeffs")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: host(*)
      end function
      integer(c_int) function ads_set_kappa(ctx, host) bind(C, name="ads_set_kappa")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         type(c_ptr), value :: host
      end function
      integer(c_int) function ads_project(ctx, f0_kind, l2norm) bind(C, name="ads_project")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: f0_kind
         real(c_double), intent(out) :: l2norm
      end function
      integer(c_int) function ads_step(ctx, dt, prec, refine, l2norm) bind(C, name="ads_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
         integer(c_int), value :: prec, refine
         real(c_double), intent(out) :: l2norm
      end function
      integer(c_int) function ads_eval_grid(ctx, res, lo, hi, vals) bind(C, name="ads_eval_grid")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), intent(in) :: res(3)
         real(c_double), intent(in) :: lo(3), hi(3)
         real(c_double), intent(out) :: vals(*)
      end function
      integer(c_int) function ads_multistep(ctx, mix, dt, l2norm) bind(C, name="ads_multistep")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: mix(4, 3)
         real(c_double), value :: dt
         real(c_double), intent(out) :: l2norm
      end function
      integer(c_int) function ads_oil_setup(ctx, cx, cy, cz, cN, cL, pumps, npumps, drains, ndrains, params) &
            bind(C, name="ads_oil_setup")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: cx(*), cy(*), cz(*), pumps(*), drains(*), params(6)
         integer(c_int), value :: cN, cL, npumps, ndrains
      end function
      integer(c_int) function ads_oil_curves(seed, cN, cL, cx, cy, cz) bind(C, name="ads_oil_curves")
         import :: c_int, c_int64_t, c_double
         integer(c_int64_t), value :: seed
         integer(c_int), value :: cN, cL
         real(c_double), intent(out) :: cx(*), cy(*), cz(*)
      end function
      integer(c_int) function ads_oil_project(ctx, l2norm) bind(C, name="ads_oil_project")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: l2norm
      end function
      integer(c_int) function ads_oil_step(ctx, dt, l2norm, drained) bind(C, name="ads_oil_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
         real(c_double), intent(out) :: l2norm, drained
      end function
      integer(c_int) function ads_form_un(ctx, subun) bind(C, name="ads_form_un")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: subun
      end function
      integer(c_int) function ads_form_rhs_sub(ctx, form, f0, direction, substep, dt, F, l2norm) &
            bind(C, name="ads_form_rhs_sub")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: form, f0, direction, substep
         real(c_double), value :: dt
         real(c_double), intent(out) :: F(*)
         real(c_double), intent(out) :: l2norm
      end function
      integer(c_int) function ads_sub_step(ctx, form, f0, mix, direction, substep, dt, l2norm) &
            bind(C, name="ads_sub_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: form, f0, direction, substep
         real(c_double), intent(in) :: mix(4)
         real(c_double), value :: dt
         real(c_double), intent(out) :: l2norm
      end function
      integer(c_int) function ads_last_drained(ctx, drained) bind(C, name="ads_last_drained")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: drained
      end function
      integer(c_int) function ads_matrix_create(U, n, p, nelem, mix, m) bind(C, name="ads_matrix_create")
         import :: c_int, c_ptr, c_double
         real(c_double), intent(in) :: U(*)
         integer(c_int), value :: n, p, nelem
         real(c_double), intent(in) :: mix(4)
         type(c_ptr), intent(out) :: m
      end function
      integer(c_int) function ads_matrix_destroy(m) bind(C, name="ads_matrix_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: m
      end function
      integer(c_int) function ads_matrix_dense(m, dense) bind(C, name="ads_matrix_dense")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: m
         real(c_double), intent(out) :: dense(*)
      end function
      integer(c_int) function ads_matrix_solve(m, rhs, eqnum) bind(C, name="ads_matrix_solve")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: m
         real(c_double), intent(inout) :: rhs(*)
         integer(c_int), value :: eqnum
      end function
      type(c_ptr) function ads_matrix_last_error(m) bind(C, name="ads_matrix_last_error")
         import :: c_ptr
         type(c_ptr), value :: m
      end function
   end interface

contains

   ! text of ads_last_error as a Fortran string
   function ads_error_text(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg
      character(kind=c_char), pointer :: chars(:)
      type(c_ptr) :: p
      integer :: n
      p = ads_last_error(ctx)
      msg = ''
      if (.not. c_associated(p)) return
      call c_f_pointer(p, chars, [4096])
      n = 0
      do while (n < 4096)
         if (chars(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate(character(len=n) :: msg)
      msg = transfer(chars(1:n), msg)
   end function

end module ads_capi
