! Replacement of the reference's susc_module (src/susc_module.f90) with the same module name, public
! procedures and argument lists: calc_chi_qw runs on the GPU through the C ABI; the two text writers
! write the reference's files (the same names, columns and number formats, which are its output
! interface).
module susc_module
  use, intrinsic :: iso_c_binding
  use parameters_module
  use adga_c
  use adga_dropin
  implicit none

contains

  subroutine calc_chi_qw(chi_qw,interm3,chi0_sum)
    complex(kind=8), intent(inout) :: chi_qw(ndim2,ndim2)
    complex(kind=8), intent(in)    :: interm3(ndim2,maxdim)
    complex(kind=8), intent(in)    :: chi0_sum(ndim2,ndim2,-iwfmax_small:iwfmax_small-1)
    integer(c_int) :: rc
    rc = c_adga_chi_qw(dropin_ctx(.true.), interm3, chi0_sum, chi_qw)
    call dropin_check(rc, 'calc_chi_qw')
  end subroutine calc_chi_qw

  ! the diagonal pairs chi(i,i,j,j), j >= i, of one (n2, n2) block as (re, im) columns
  subroutine diag_pairs(chi, vals, nv)
    complex(kind=8), intent(in) :: chi(ndim2,ndim2)
    real(kind=8), intent(out) :: vals(:)
    integer, intent(out) :: nv
    integer :: i, j
    nv = 0
    do i = 1, ndim
      do j = i, ndim
        vals(nv+1) = real(chi((i-1)*ndim+i,(j-1)*ndim+j))
        vals(nv+2) = aimag(chi((i-1)*ndim+i,(j-1)*ndim+j))
        nv = nv + 2
      enddo
    enddo
  end subroutine diag_pairs

  subroutine output_chi_qw(chi_qw,qw,filename_output)
    complex(kind=8),intent(in)  :: chi_qw(ndim2,ndim2,nqp*(2*iwbmax_small+1))
    integer,intent(in)          :: qw(2,nqp*(2*iwbmax+1))
    character(len=*),intent(in) :: filename_output
    real(kind=8) :: vals(ndim*(ndim+1))
    integer :: p, nv, iwb, iq
    open(unit=10, file=trim(output_dir)//filename_output)
    write(10,'(A)') '#iwbmax, nqp, ndim, beta, mu'
    write(10,'(I5,2X,I5,2X,I5,2X,E14.7E2,2X,E14.7E2)') iwbmax_small, nqp, ndim, beta, mu
    write(10,'(A)') '#iwb, wb, iq, qx, qy, qz, (real(chi_qw(wb,q,i,i,j,j)), imag(chi_qw(wb,q,i,i,j,j))) [j=i,ndim] [i=1,ndim]'
    do p = 1, nqp*(2*iwbmax_small+1)
      iwb = qw(1,p)
      iq = qw(2,p)
      call diag_pairs(chi_qw(:,:,p), vals, nv)
      write(10,'(I8,E14.7E2,I8,9999E18.7E2)') iwb, iwb_data(iwb), iq, k_data(:,q_data(iq)), vals(1:nv)
      if (mod(p, nqp) == 0) write(10,*) ' '   ! a blank line after each omega block
    enddo
    if (ounit .ge. 1 .and. (verbose .and. (index(verbstr,"Output") .ne. 0))) &
      write(ounit,*) 'nonlocal susceptibility written to ', trim(filename_output)
    close(10)
  end subroutine output_chi_qw

  subroutine output_chi_loc(chi_w,filename_output)
    complex(kind=8),intent(in)  :: chi_w(ndim2,ndim2,2*iwbmax_small+1)
    character(len=*),intent(in) :: filename_output
    real(kind=8) :: vals(ndim*(ndim+1))
    integer :: p, nv, iwb
    open(unit=10, file=trim(output_dir)//filename_output)
    write(10,'(A)') '#iwbmax, nqp, ndim, beta, mu'
    write(10,'(I5,2X,I5,2X,I5,2X,E14.7E2,2X,E14.7E2)') iwbmax_small, 0, ndim, beta, mu
    write(10,'(A)') '#iwb, wb, (real(chi_loc(wb,i,i,j,j)), imag(chi_loc(wb,i,i,j,j))) [j=i,ndim] [i=1,ndim]'
    do p = 1, 2*iwbmax_small+1
      iwb = p - iwbmax_small - 1
      call diag_pairs(chi_w(:,:,p), vals, nv)
      write(10,'(I8,9999E18.7E2)') iwb, iwb_data(iwb), vals(1:nv)
    enddo
    close(10)
    if (ounit .ge. 1 .and. (verbose .and. (index(verbstr,"Output") .ne. 0))) &
      write(ounit,*) 'local susceptibility written to ', trim(filename_output)
  end subroutine output_chi_loc

end module susc_module
