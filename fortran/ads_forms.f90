! ads_forms.f90 -- the table that binds a problem's right-hand-side callback
! (the RHS_fun argument of Step, MultiStep, Sub_Step and Form3DRHS,
! src/Interfaces.F90:52-81) to a device form.  The callback itself cannot run
! on the GPU: a problem registers it once (ADSS's RegisterHeatForm,
! RegisterProjectionForm, RegisterOilForm) and the entry points look it up
! by its address (c_funloc).
module ads_forms
   use iso_c_binding
   use Interfaces, only: RHS_fun_int
   use ads_capi
   implicit none
   private
   public :: form_entry, add_form, find_form, form_at_t0, report
   public :: FORM_HEAT, FORM_PROJECTION, FORM_OIL

   integer, parameter :: FORM_HEAT = 1, FORM_PROJECTION = 2, FORM_OIL = 3, MAXFORMS = 16

   type form_entry
      type(c_funptr) :: fptr = c_null_funptr
      integer :: kind = 0
      integer(c_int) :: f0 = ADS_F0_CONST
      real(c_double) :: dt = 0.d0
      integer(c_int) :: prec = ADS_PREC_FP64
      integer(c_int) :: refine = 2
      ! the problem's own time variable (input_data's t) when registered with
      ! it: the heat integrand takes its t == 0 branch exactly when t > 0 fails
      type(c_ptr) :: tptr = c_null_ptr
   end type form_entry

   type(form_entry), save :: forms(MAXFORMS)
   integer, save :: nforms = 0

contains

   subroutine report(ctx, code, what)
      type(c_ptr), intent(in) :: ctx
      integer(c_int), intent(in) :: code
      character(len=*), intent(in) :: what
      if (code /= ADS_OK) write(*, '(a,a,i0,a,a)') what, ': ADS error ', code, ': ', ads_error_text(ctx)
   end subroutine

   subroutine add_form(RHS_fun, kind, f0, dt, prec, refine, tptr, mierr)
      procedure(RHS_fun_int) :: RHS_fun
      integer, intent(in) :: kind
      integer(c_int), intent(in) :: f0, prec, refine
      real(c_double), intent(in) :: dt
      type(c_ptr), intent(in) :: tptr
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
      forms(i) = form_entry(fp, kind, f0, dt, prec, refine, tptr)
      mierr = 0
   end subroutine

   ! The entry bound to RHS_fun; .false. (and a message naming the caller)
   ! when the procedure was never registered.
   logical function find_form(RHS_fun, what, f) result(found)
      procedure(RHS_fun_int) :: RHS_fun
      character(len=*), intent(in) :: what
      type(form_entry), intent(out) :: f
      type(c_funptr) :: fp
      integer :: i
      fp = c_funloc(RHS_fun)
      found = .false.
      do i = 1, nforms
         if (c_associated(forms(i)%fptr, fp)) then
            f = forms(i)
            found = .true.
            return
         end if
      end do
      write(*, '(a,a)') what, ': RHS_fun is not bound to a device form (RegisterHeatForm/RegisterOilForm/...)'
   end function

   ! Whether the form takes its t = 0 branch (the initial-state projection):
   ! by the problem's own t when registered with it (the integrand's test t >
   ! 0), else by iter == 0 (and, without an iteration, by ads_data%t == 0).
   logical function form_at_t0(f, iter, tdata) result(t0)
      type(form_entry), intent(in) :: f
      integer(kind=4), intent(in) :: iter
      real(kind=8), intent(in) :: tdata
      real(c_double), pointer :: tp
      if (c_associated(f%tptr)) then
         call c_f_pointer(f%tptr, tp)
         t0 = .not. (tp > 0.0)
      else if (iter >= 0) then
         t0 = iter == 0
      else
         t0 = .not. (tdata > 0.d0)
      end if
   end function

end module ads_forms
