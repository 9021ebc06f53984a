! sparse.f90 -- drop-in for module sparse (src/sparse.F90): the type that
! ComputeMatrix fills and SolveOneDirection reads.  The reference keeps the
! 1-D matrix as a sorted linked list and hands it to MUMPS; here the
! matrix object of the C ABI (ads_matrix_create) holds the band of the same
! matrix and its banded LU with partial pivoting, and SolveOneDirection runs
! the substitution sweeps on the GPU.
module sparse
   use iso_c_binding
   use ads_capi
   implicit none
   private
   public :: sparse_matrix, clear_matrix, matrix_error_text

   type sparse_matrix
      ! the factored matrix (ads_matrix) and its size: (n+1) x (n+1), degree p
      type(c_ptr) :: handle = c_null_ptr
      integer(kind=4) :: n = -1, p = 0
   end type sparse_matrix

contains

   ! frees the matrix object and the type (src/sparse.F90:189-200)
   subroutine clear_matrix(matrix)
      type(sparse_matrix), pointer, intent(inout) :: matrix
      integer(c_int) :: rc
      if (.not. associated(matrix)) return
      if (c_associated(matrix%handle)) rc = ads_matrix_destroy(matrix%handle)
      deallocate(matrix)
      nullify(matrix)
   end subroutine clear_matrix

   ! text of the matrix object's last error
   function matrix_error_text(matrix) result(msg)
      type(sparse_matrix), intent(in) :: matrix
      character(len=:), allocatable :: msg
      character(kind=c_char), pointer :: chars(:)
      type(c_ptr) :: p
      integer :: n
      msg = ''
      p = ads_matrix_last_error(matrix%handle)
      if (.not. c_associated(p)) return
      call c_f_pointer(p, chars, [4096])
      n = 0
      do while (n < 4096)
         if (chars(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate(character(len=n) :: msg)
      msg = transfer(chars(1:n), msg)
   end function matrix_error_text

end module sparse
