! math.f90 -- drop-in for module math (src/math.F90): the helper functions
! the reference's problem sources import (problems/*/input_data.F90: lerp,
! falloff, bump3d, norm2), with the same names, arguments and formulas.
module math
   implicit none
   private
   public :: PI, lerp, falloff, bump, bump01, bump3d, norm2

   real(kind=8), parameter :: PI = 4.d0 * atan(1.d0)

contains

   ! (1 - t) x + t y: x at t = 0, y at t = 1 (src/math.F90:44-51)
   pure function lerp(t, x, y) result(val)
      real(kind=8), intent(in) :: t, x, y
      real(kind=8) :: val
      val = (1.d0 - t) * x + t * y
   end function lerp

   ! 1 below r, 0 above Rr, and (h^2 - 1)^2 with h = (t - r)/(Rr - r) between:
   ! C^1, with value and slope vanishing at both ends (src/math.F90:74-89)
   pure function falloff(r, Rr, t) result(fval)
      real(kind=8), intent(in) :: r, Rr, t
      real(kind=8) :: fval, h
      if (t < r) then
         fval = 1.d0
      else if (t > Rr) then
         fval = 0.d0
      else
         h = (t - r) / (Rr - r)
         fval = ((h - 1.d0) * (h + 1.d0))**2.d0
      end if
   end function falloff

   ! (x + 1)^2 (x - 1)^2 on (-1, 1), 0 elsewhere (src/math.F90:108-119)
   pure function bump(x) result(val)
      real(kind=8), intent(in) :: x
      real(kind=8) :: val
      val = 0.d0
      if (x > -1.d0 .and. x < 1.d0) val = (x + 1.d0)**2.d0 * (x - 1.d0)**2.d0
   end function bump

   ! bump scaled to (0, 1) (src/math.F90:139-146)
   pure function bump01(x) result(val)
      real(kind=8), intent(in) :: x
      real(kind=8) :: val
      val = bump(2.d0 * (x - 0.5d0))
   end function bump01

   ! falloff(r/2, Rr/2, |(x, y, z) - (1/2, 1/2, 1/2)|) (src/math.F90:173-182)
   pure function bump3d(r, Rr, x, y, z) result(val)
      real(kind=8), intent(in) :: r, Rr, x, y, z
      real(kind=8) :: val
      val = falloff(r / 2.d0, Rr / 2.d0, norm2((/x, y, z/) - 0.5d0))
   end function bump3d

   ! Euclidean norm as sqrt(dot_product(x, x)) (src/math.F90:198-204)
   pure function norm2(x) result(val)
      real(kind=8), intent(in) :: x(:)
      real(kind=8) :: val
      val = sqrt(dot_product(x, x))
   end function norm2

end module math
