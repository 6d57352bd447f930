! Drop-in replacement for the reference's lapack_module (src/lapack_module.f90:28-104):
! same module name, same public generic `inverse_matrix` with the same specific
! signatures, so main.F90 and the other modules compile against it unchanged.
! The complex inverse runs on the GPU (LU with partial pivoting and n solves,
! libadga_amd.so) up to n = 4096 (the batched LU's largest size); a larger matrix
! is inverted on the host with LAPACK's zgetrf + zgetri, which the reference's own
! build links anyway.  Errors keep the reference's contract: ierr /= 0 and a
! message in erstr ("ERROR in ZGETRF" for an exactly singular matrix).
module lapack_module
  use, intrinsic :: iso_c_binding
  use adga_c, only: c_adga_inverse_matrix_z, ADGA_OK, ADGA_E_SINGULAR
  implicit none
  private
  public inverse_matrix

  integer, parameter :: gpu_max_n = 4096   ! ADGA_INVERSE_MAX_N of include/adga_amd.h

  interface inverse_matrix
    module procedure inverse_matrix_d, inverse_matrix_z
  end interface inverse_matrix

contains

  subroutine inverse_matrix_z(M, erstr, ierr)
    double complex, intent(inout)   :: M(:,:)
    integer, intent(out)            :: ierr
    character(len=200), intent(out) :: erstr
    complex(c_double_complex), allocatable :: work(:,:)
    integer(c_int) :: info, rc
    integer :: n

    n = size(M, 1)
    erstr = ''
    if (size(M, 2) /= n) then
      ierr = -1
      erstr = "ERROR: inverse_matrix needs a square matrix"
      return
    endif
    if (n > gpu_max_n) then
      call inverse_host(M, erstr, ierr)
      return
    endif
    allocate(work(n, n))
    work = M                                 ! contiguous copy (M may be a strided section)
    rc = c_adga_inverse_matrix_z(int(n, c_int), work, info)
    ierr = int(info)
    if (rc == ADGA_E_SINGULAR) then
      erstr = "ERROR in ZGETRF"
      if (ierr == 0) ierr = 1
    else if (rc /= ADGA_OK) then
      erstr = "ERROR in the GPU inverse (adga_inverse_matrix_z)"
      if (ierr == 0) ierr = int(rc)
    else
      M = work
    endif
    deallocate(work)
  end subroutine inverse_matrix_z

  ! zgetrf + workspace query + zgetri on the host, for sizes above the GPU path's limit
  subroutine inverse_host(M, erstr, ierr)
    double complex, intent(inout)   :: M(:,:)
    integer, intent(out)            :: ierr
    character(len=200), intent(out) :: erstr
    double complex, allocatable :: a(:,:), work(:)
    double complex :: wq(1)
    integer, allocatable :: ipiv(:)
    integer :: n, lwork

    n = size(M, 1)
    allocate(a(n, n), ipiv(n))
    a = M
    call zgetrf(n, n, a, n, ipiv, ierr)
    if (ierr /= 0) then
      erstr = "ERROR in ZGETRF"
      deallocate(a, ipiv)
      return
    endif
    call zgetri(n, a, n, ipiv, wq, -1, ierr)
    lwork = max(1, int(dble(wq(1))))
    allocate(work(lwork))
    call zgetri(n, a, n, ipiv, work, lwork, ierr)
    if (ierr /= 0) then
      erstr = "ERROR in ZGETRI"
    else
      M = a
    endif
    deallocate(a, ipiv, work)
  end subroutine inverse_host

  subroutine inverse_matrix_d(M, erstr, ierr)
    double precision, intent(inout) :: M(:,:)
    integer, intent(out)            :: ierr
    character(len=200), intent(out) :: erstr
    double complex, allocatable :: z(:,:)

    allocate(z(size(M, 1), size(M, 2)))
    z = cmplx(M, 0d0, kind=8)
    call inverse_matrix_z(z, erstr, ierr)
    if (ierr == 0) M = dble(z)
    if (erstr == "ERROR in ZGETRF") erstr = "ERROR in DGETRF"
    deallocate(z)
  end subroutine inverse_matrix_d

end module lapack_module
