! projection_engine.f90 -- drop-in for module projection_engine
! (src/projection_engine.F90): FormUn, Form3DRHS, MKBBT/ComputeMatrix and
! global2local with the reference's argument lists.  FormUn and Form3DRHS run
! on the GPU through the C ABI: the point values and gradients are formed
! inside the assembly kernels, so FormUn only fixes which coefficients they
! come from, and Form3DRHS evaluates the device form bound to RHS_fun
! (ads_forms) and leaves the right-hand side in ads_data%F.
module projection_engine
   use iso_c_binding
   use Setup, only: ADS_setup, ADS_compute_data
   use Interfaces, only: RHS_fun_int
   use ads_capi
   use ads_forms
   use sparse, only: sparse_matrix, matrix_error_text
   implicit none
   private
   public :: FormUn, Form3DRHS, MKBBT, ComputeMatrix, global2local

contains

   ! FormUn (:306-441).  subun = 1 keeps the current coefficients as U_n,
   ! whose point values every later right-hand side of the step uses (the
   ! reference's Un); subun = 2, 3 (Un13, Un23) keep U_n, and the gradients
   ! (dUn) always come from the latest coefficients.
   subroutine FormUn(subun, ads, ads_data)
      integer(kind=4), intent(in) :: subun
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      integer(c_int) :: rc
      if (subun < 1 .or. subun > 3) then
         write(*, '(a)') 'wrong substep'
         return
      end if
      rc = ads_form_un(ads%handle, int(subun, c_int))
      call report(ads%handle, rc, 'FormUn')
   end subroutine FormUn

   ! Form3DRHS (:154-284): ads_data%F (x fastest, F(s1, s2*s3)) receives the
   ! right-hand side of the device form bound to RHS_fun for the current
   ! coefficients, l2norm its l2 sum (sum of J W v u^2, or J W v f0^2 in the
   ! t = 0 branch).  The heat form's t = 0 branch (the projection of its
   ! initial state) is taken when its registered time variable is not
   ! positive, else when ads_data%t is not.
   subroutine Form3DRHS(ads, ads_data, direction, substep, RHS_fun, l2norm)
      type(ADS_setup), intent(in) :: ads
      type(ADS_compute_data), intent(inout) :: ads_data
      integer(kind=4), intent(in) :: direction, substep
      procedure(RHS_fun_int) :: RHS_fun
      real(kind=8), intent(out) :: l2norm
      type(form_entry) :: f
      integer(c_int) :: rc, form
      real(c_double) :: l2
      l2norm = 0.d0
      if (.not. find_form(RHS_fun, 'Form3DRHS', f)) return
      form = ADS_FORM_HEAT
      if (f%kind == FORM_OIL) form = ADS_FORM_OIL
      if (f%kind == FORM_OIL .and. .not. (ads_data%t > 0.d0)) form = ADS_FORM_OIL_PROJECT
      if (f%kind == FORM_PROJECTION) form = ADS_FORM_PROJECT
      if (f%kind == FORM_HEAT .and. form_at_t0(f, -1, ads_data%t)) form = ADS_FORM_PROJECT
      rc = ads_form_rhs_sub(ads%handle, form, f%f0, int(direction, c_int), int(substep, c_int), f%dt, ads_data%F, l2)
      call report(ads%handle, rc, 'Form3DRHS')
      if (rc == ADS_OK) l2norm = l2
   end subroutine Form3DRHS

   ! MKBBT (:63-132): the 1-D matrix mix(1) M + mix(2) K + mix(3) B + mix(4)
   ! B^T of the degree-p B-splines on the knot vector U(0:n+p+1), assembled
   ! by Gauss quadrature over the nelem elements, into a new sparse_matrix
   ! (factored for SolveOneDirection).  sprsmtrx is null on failure.
   subroutine MKBBT(U, p, n, nelem, mix, sprsmtrx)
      integer(kind=4), intent(in) :: n, p, nelem
      real(kind=8), dimension(0:n + p + 1), intent(in) :: U
      real(kind=8), dimension(4), intent(in) :: mix
      type(sparse_matrix), pointer, intent(out) :: sprsmtrx
      integer(c_int) :: rc
      allocate(sprsmtrx)
      sprsmtrx%n = n
      sprsmtrx%p = p
      rc = ads_matrix_create(U, int(n, c_int), int(p, c_int), int(nelem, c_int), mix, sprsmtrx%handle)
      if (rc /= ADS_OK) then
         write(*, '(a,i0,a,a)') 'MKBBT: ADS error ', rc, ': ', matrix_error_text(sprsmtrx)
         deallocate(sprsmtrx)
         nullify(sprsmtrx)
      end if
   end subroutine MKBBT

   ! ComputeMatrix (:531-544)
   subroutine ComputeMatrix(U, p, n, nelem, mix, sprsmtrx)
      integer(kind=4), intent(in) :: n, p, nelem
      real(kind=8), dimension(0:n + p + 1), intent(in) :: U
      real(kind=8), dimension(4), intent(in) :: mix
      type(sparse_matrix), pointer, intent(out) :: sprsmtrx
      call MKBBT(U, p, n, nelem, mix, sprsmtrx)
   end subroutine ComputeMatrix

   ! global2local (:495-507): the 0-based (x, y, z) of the x-fastest index
   ! ind in a cube of (n(1)+1) x (n(2)+1) x (n(3)+1) coefficients.
   subroutine global2local(ind, n, x, y, z)
      integer(kind=4), intent(in) :: ind
      integer(kind=4), dimension(3), intent(in) :: n
      integer(kind=4), intent(out) :: x, y, z
      integer(kind=4) :: plane, rest
      plane = (n(1) + 1) * (n(2) + 1)
      z = ind / plane
      rest = ind - z * plane
      y = rest / (n(1) + 1)
      x = rest - y * (n(1) + 1)
   end subroutine global2local

end module projection_engine
