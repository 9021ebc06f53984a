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

   ! F holds the coefficients (x fastest, F(s1, s2*s3)) after every Step, as
   ! in the reference, and the right-hand side after Form3DRHS.  The other
   ! components keep the reference's names and ranks so that problem code
   ! naming them compiles, but stay unallocated: the transposed copies (F2,
   ! F3, F*_out), the neighbour buffer R and the quadrature-point arrays (Un,
   ! Un13, Un23, dUn) are never formed, because FormUn is fused into the
   ! assembly kernels and the transposes happen on the device.
   type ADS_compute_data
      real(kind=8), allocatable, dimension(:,:) :: F, F2, F3
      real(kind=8), allocatable, dimension(:,:) :: F_out, F2_out, F3_out
      real(kind=8), allocatable, dimension(:,:,:,:) :: R
      real(kind=8) :: t = 0.d0
      real(kind=8), allocatable, dimension(:,:,:,:,:,:) :: Un
      real(kind=8), allocatable, dimension(:,:,:,:,:,:) :: Un13
      real(kind=8), allocatable, dimension(:,:,:,:,:,:) :: Un23
      real(kind=8), allocatable, dimension(:,:,:,:,:,:,:) :: dUn
   end type ADS_compute_data

end module Setup
