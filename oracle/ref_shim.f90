! ref_shim.f90 -- bind(C) entry points onto the reference's own MPI-free
! leaf modules (knot_vector, gauss, basis, reorderRHS, math), compiled with flang
! from the sources where they lie.  TEST INFRASTRUCTURE ONLY: the result,
! oracle/_ref/libadsref.so, pins the oracle's restatement bitwise; the
! product library never links it.
module ads_ref_shim
   use iso_c_binding
   use knot_vector, only: FillOpenKnot, kv_count => CountSpans
   use gauss, only: GaussRule
   use basis, only: BasisData, DersBasisFuns, FindSpan, EvalSpline
   use reorderRHS, only: ReorderRHSForX, ReorderRHSForY, ReorderRHSForZ
   use math, only: falloff, lerp, bump3d
   implicit none
contains

   function ref_falloff(r, Rr, t) result(v) bind(C, name="ref_falloff")
      real(c_double), value :: r, Rr, t
      real(c_double) :: v
      v = falloff(r, Rr, t)
   end function

   function ref_lerp(t, x, y) result(v) bind(C, name="ref_lerp")
      real(c_double), value :: t, x, y
      real(c_double) :: v
      v = lerp(t, x, y)
   end function

   function ref_bump3d(r, Rr, x, y, z) result(v) bind(C, name="ref_bump3d")
      real(c_double), value :: r, Rr, x, y, z
      real(c_double) :: v
      v = bump3d(r, Rr, x, y, z)
   end function

   subroutine ref_fill_open_knot(n, p, U) bind(C, name="ref_fill_open_knot")
      integer(c_int), value :: n, p
      real(c_double), intent(out) :: U(n + p + 2)
      call FillOpenKnot(n, p, U)
   end subroutine

   function ref_count_spans(n, p, U) result(ne) bind(C, name="ref_count_spans")
      integer(c_int), value :: n, p
      real(c_double), intent(in) :: U(0:n + p + 1)
      integer(c_int) :: ne
      ne = kv_count(n, p, U)
   end function

   subroutine ref_gauss_rule(q, X, W) bind(C, name="ref_gauss_rule")
      integer(c_int), value :: q
      real(c_double), intent(out) :: X(0:q - 1), W(0:q - 1)
      call GaussRule(q, X, W)
   end subroutine

   subroutine ref_basis_data(p, m, U, d, q, r, O, J, W, X, NN) bind(C, name="ref_basis_data")
      integer(c_int), value :: p, m, d, q, r
      real(c_double), intent(in) :: U(0:m)
      integer(c_int), intent(out) :: O(r)
      real(c_double), intent(out) :: J(r), W(q), X(q, r), NN(0:d, 0:p, q, r)
      call BasisData(p, m, U, d, q, r, O, J, W, X, NN)
   end subroutine

   subroutine ref_ders_basis_funs(i, uu, p, d, U, ders) bind(C, name="ref_ders_basis_funs")
      integer(c_int), value :: i, p, d
      real(c_double), value :: uu
      real(c_double), intent(in) :: U(0:i + p)
      real(c_double), intent(out) :: ders(0:p, 0:d)
      call DersBasisFuns(i, uu, p, d, U, ders)
   end subroutine

   function ref_find_span(n, p, uu, U) result(span) bind(C, name="ref_find_span")
      integer(c_int), value :: n, p
      real(c_double), value :: uu
      real(c_double), intent(in) :: U(0:n + p + 1)
      integer(c_int) :: span
      span = FindSpan(n, p, uu, U)
   end function

   function ref_eval_spline(nx, px, ny, py, nz, pz, Ux, Uy, Uz, coeffs, x, y, z) result(val) &
         bind(C, name="ref_eval_spline")
      integer(c_int), value :: nx, px, ny, py, nz, pz
      real(c_double), intent(in) :: Ux(0:nx + px + 1), Uy(0:ny + py + 1), Uz(0:nz + pz + 1)
      real(c_double), intent(in) :: coeffs(0:nx, 0:ny, 0:nz)
      real(c_double), value :: x, y, z
      real(c_double) :: val
      val = EvalSpline(0, Ux, px, nx, nx - px + 1, Uy, py, ny, ny - py + 1, &
                       Uz, pz, nz, nz - pz + 1, coeffs, x, y, z)
   end function

   subroutine ref_reorder(which, s, F, F2) bind(C, name="ref_reorder")
      integer(c_int), value :: which
      integer(c_int), intent(in) :: s(3)
      real(c_double), intent(in) :: F(*)
      real(c_double), intent(out) :: F2(*)
      integer(kind=4) :: ibeg(3), iend(3)
      ibeg = 1
      iend = s
      select case (which)
      case (1)
         call ReorderRHSForY(ibeg, iend, F, F2)
      case (2)
         call ReorderRHSForZ(ibeg, iend, F, F2)
      case default
         call ReorderRHSForX(ibeg, iend, F, F2)
      end select
   end subroutine

end module ads_ref_shim
