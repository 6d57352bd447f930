! Replacement of the reference's one_particle_quant_module (src/one_particle_quant_module.f90) with the
! same module name, public procedures and argument lists, so main.F90 and eom_module compile against it
! unchanged; every procedure runs on the GPU through the C ABI (include/adga_amd.h).  Compiled against
! the reference's parameters_module, whose globals (hk, siw, giw, dc, mu, beta, ...) are the inputs.
!
! accumulate_chi0 relies on the calling loop of main.F90:647-653 (for each iwf, ik = 1 .. nkp with
! ikq = kq_ind(ik, iq)): at ik == 1 it adds the whole k sum for that (q, nu) and later ik add nothing.
! The q follows from ikq = kq_ind(1, iq), the index of k_1 - q with k_1 = 0.  The bubble of all nu of
! the box is formed on the GPU at the first call of a (q, omega) and kept for the next nu.
module one_particle_quant_module
  use, intrinsic :: iso_c_binding
  use lapack_module
  use parameters_module
  use adga_c
  use adga_dropin
  use mpi_org, only: mpi_wrank, master, mpi_stop
  implicit none

  integer, save, private :: bub_q = -1, bub_w = huge(0), loc_w = huge(0)
  complex(kind=8), allocatable, save, private :: bub(:,:,:), loc(:,:), loc_inv(:,:)

contains

subroutine get_giw()
  integer(c_int) :: rc
  rc = c_adga_get_giw(dropin_ctx(.false.), giw)
  call dropin_check(rc, 'get_giw')
  loc_w = huge(0)   ! the local bubble follows giw
end subroutine get_giw

subroutine get_gkiw(ikq, iwf, iwb, gkiw)
  integer :: iwf, iwb, ikq
  complex(kind=8), intent(out) :: gkiw(ndim,ndim)
  integer(c_int) :: rc
  rc = c_adga_get_gkiw(dropin_ctx(.false.), int(ikq, c_int), int(iwf - iwb, c_int), c_null_ptr, gkiw)
  call dropin_check(rc, 'get_gkiw')
end subroutine get_gkiw

subroutine get_gkiw_dga(ik, iwf, skiw, gkiw)
  integer, intent(in) :: iwf, ik
  complex(kind=8), intent(in) :: skiw(ndim,ndim)
  complex(kind=8), intent(out) :: gkiw(ndim,ndim)
  complex(kind=8), target :: s(ndim,ndim)
  integer(c_int) :: rc
  s = skiw
  rc = c_adga_get_gkiw(dropin_ctx(.false.), int(ik, c_int), int(iwf, c_int), c_loc(s), gkiw)
  call dropin_check(rc, 'get_gkiw_dga')
end subroutine get_gkiw_dga

subroutine get_sigma_g_loc(sigma_sum, sigma_loc, gloc)
  complex(kind=8),intent(in) :: sigma_sum(ndim,ndim,-iwfmax_small:iwfmax_small-1,nkp)
  complex(kind=8), intent(out) :: sigma_loc(ndim, ndim,-iwfmax_small:iwfmax_small-1)
  complex(kind=8), intent (out) :: gloc(-iwmax:iwmax-1,ndim,ndim)
  integer(c_int) :: rc
  rc = c_adga_get_sigma_g_loc(dropin_ctx(.true.), sigma_sum, sigma_loc, gloc)
  call dropin_check(rc, 'get_sigma_g_loc')
end subroutine get_sigma_g_loc

! diagonals of chi0^w and its inverse for every nu of the box at omega = iwb
subroutine local_bubble(iwb)
  integer, intent(in) :: iwb
  integer(c_int) :: rc
  if (iwb == loc_w) return
  if (.not. allocated(loc)) allocate(loc(ndim2, iwstart:iwstop), loc_inv(ndim2, iwstart:iwstop))
  rc = c_adga_chi0_loc(dropin_ctx(.true.), int(iwb, c_int), loc, loc_inv)
  call dropin_check(rc, 'get_chi0_loc')
  loc_w = iwb
end subroutine local_bubble

subroutine get_chi0_loc(iwf, iwb, chi0_loc)
  integer :: iwf, iwb
  complex(kind=8), intent(out) :: chi0_loc(ndim*ndim,ndim*ndim)
  integer :: i1
  call local_bubble(iwb)
  chi0_loc = 0.d0
  do i1 = 1, ndim*ndim
    chi0_loc(i1,i1) = loc(i1,iwf)
  enddo
end subroutine get_chi0_loc

subroutine get_chi0_loc_inv(iwf, iwb, chi0_loc)
  integer :: iwf, iwb
  complex(kind=8), intent(out) :: chi0_loc(ndim*ndim,ndim*ndim)
  integer :: i1
  call local_bubble(iwb)
  chi0_loc = 0.d0
  do i1 = 1, ndim*ndim
    chi0_loc(i1,i1) = loc_inv(i1,iwf)
  enddo
end subroutine get_chi0_loc_inv

subroutine accumulate_chi0(ik, ikq, iwf, iwb, chi0)
  integer :: iwf, iwb, ik, ikq
  complex(kind=8), intent(inout) :: chi0(ndim*ndim,ndim*ndim)
  integer :: kx, ky, kz, qk
  integer(c_int) :: rc
  if (ik /= 1) return
  ! ikq = index of -q: negate its grid coordinates (k ordering of kq_tools' k_index_3)
  kz = mod(ikq - 1, nkpz)
  ky = mod((ikq - 1)/nkpz, nkpy)
  kx = (ikq - 1)/(nkpz*nkpy)
  qk = 1 + mod(nkpz - kz, nkpz) + mod(nkpy - ky, nkpy)*nkpz + mod(nkpx - kx, nkpx)*nkpy*nkpz
  if (qk /= bub_q .or. iwb /= bub_w) then
    if (.not. allocated(bub)) allocate(bub(ndim2, ndim2, iwstart:iwstop))
    rc = c_adga_bubble_q(dropin_ctx(.true.), int(iwb, c_int), int(qk, c_int), bub)
    call dropin_check(rc, 'accumulate_chi0')
    bub_q = qk
    bub_w = iwb
  endif
  chi0 = chi0 + dble(nkp)*bub(:,:,iwf)   ! the caller divides the k sum by nkp
end subroutine accumulate_chi0

subroutine get_ndmft()
  integer :: i
  complex(c_double_complex) :: nd(ndim)
  integer(c_int) :: rc
  rc = c_adga_get_ndmft(dropin_ctx(.false.), nd)
  call dropin_check(rc, 'get_ndmft')
  n_dmft = nd
  if ((verbose .and. (index(verbstr,"Dmft") .ne. 0)) .and. mpi_wrank .eq. master) then
    open(56, file=trim(output_dir)//"n_dmft.dat", status='unknown')
    write(56,*) '# n_dmft(i,i) [i=1,ndim]'
    write(56,'(100F12.6)') (real(n_dmft(i)), i=1,ndim)
    close(56)
  endif
end subroutine get_ndmft

subroutine get_nfock()
  integer :: ik, i
  integer(c_int) :: rc
  rc = c_adga_get_nfock(dropin_ctx(.false.), n_fock)
  call dropin_check(rc, 'get_nfock')
  if (mpi_wrank .eq. master .and. text_output) then
    open(110, file=trim(output_dir)//"n_fock.dat", status='unknown')
    write(110,*) '# ik, kx, ky, kz, n_fock(k,i,i) [i=1,ndim]'
    do ik = 1, nkp
      write(110,'(I8, 100F12.6)') ik, k_data(:,ik), (real(n_fock(ik,i,i)), i=1,ndim)
    enddo
    close(110)
  endif
end subroutine get_nfock

subroutine get_ndga(sigma_sum)
  complex(kind=8), intent(in) :: sigma_sum(ndim,ndim,-iwfmax_small:iwfmax_small-1, nkp_eom)
  integer :: ik, i
  integer(c_int) :: rc
  rc = c_adga_get_ndga(dropin_ctx(.true.), sigma_sum, n_dga_k, n_dga)
  call dropin_check(rc, 'get_ndga')
  if (mpi_wrank .eq. master .and. text_output) then
    open(110, file=trim(output_dir)//"n_dga_k.dat", status='unknown')
    write(110,*) '# ik, kx, ky, kz, n_dga(k,i,i) [i=1,ndim]'
    do ik = 1, nkp_eom
      write(110,'(I8,100F12.6)') ik, k_data(:,k_data_eom(ik)), (real(n_dga_k(ik,i,i)), i=1,ndim)
    enddo
    close(110)
    if (.not. k_path_eom) then
      open(111, file=trim(output_dir)//"n_dga.dat", status='unknown')
      write(111,*) '# n_dga(i,i) [i=1,ndim]'
      write(111,'(100F12.6)') (real(n_dga(i)), i=1,ndim)
      close(111)
    endif
  endif
end subroutine get_ndga

end module one_particle_quant_module
