! Setup.f90 -- drop-in for module Setup (src/Setup.F90:29-140): the solver
! configuration and state types with the reference's public component names,
! on one rank, plus the handle of the GPU context behind them.
module Setup
   use iso_c_binding, only: c_ptr, c_null_ptr
   implicit none

   type ADS_setup
      integer(kind=4), dimension(3) :: n, p, nelem
      real(kind=8), allocatable, dimension(:) :: Ux, Uy, Uz
      integer(kind=4), allocatable, dimension(:) :: dimensionsX, dimensionsY, dimensionsZ
      integer(kind=4), allocatable, dimension(:) :: shiftsX, shiftsY, shiftsZ
      integer(kind=4), allocatable, dimension(:) :: IPIVx, IPIVy, IPIVz
      integer(kind=4), dimension(3) :: nrcpp, ibeg, iend, s
      integer(kind=4), dimension(3) :: ibegsx, iendsx, ibegsy, iendsy, ibegsz, iendsz
      integer(kind=4), dimension(3) :: mine, maxe, m, ng
      integer(kind=4), allocatable, dimension(:) :: Ox, Oy, Oz
      real(kind=8), allocatable, dimension(:) :: Jx, Jy, Jz
      real(kind=8), allocatable, dimension(:,:) :: Xx, Xy, Xz
      real(kind=8), allocatable, dimension(:,:,:,:) :: NNx, NNy, NNz
      real(kind=8), allocatable, dimension(:) :: Wx, Wy, Wz
      integer(kind=4), dimension(3) :: lnelem
      real(kind=8) :: tau = 0.d0
      ! the GPU context (ads_init) that owns the device state
      type(c_ptr) :: handle = c_null_ptr
   end type ADS_setup

   ! The reference's Un/Un13/Un23/dUn quadrature-point arrays (and R, F2, F3,
   ! F_out...) are never formed: FormUn is fused into the assembly kernel and
   ! the transposes happen on the device.  F holds the coefficients (x fastest,
   ! F(s1, s2*s3)) after every Step, as in the reference.
   type ADS_compute_data
      real(kind=8), allocatable, dimension(:,:) :: F
      real(kind=8) :: t = 0.d0
   end type ADS_compute_data

end module Setup
