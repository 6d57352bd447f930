! Replacement of the reference's eom_module (src/eom_module.f90) with the same module name, public
! procedures and argument lists: calc_eom_static, calc_eom_dmft, calc_eom_dynamic and add_siw_dmft run
! on the GPU through the C ABI; output_eom writes the reference's text files (the same names, columns
! and number formats, which are its output interface).  k - q comes from the context's q and EOM k
! lists, i.e. the table index_kq_eom builds (the kq_ind_eom argument is not read).
module eom_module
  use, intrinsic :: iso_c_binding
  use parameters_module
  use one_particle_quant_module
  use adga_c
  use adga_dropin
  use mpi_org, only: mpi_stop
  implicit none

contains

  ! V(q) for the GPU when do_vq, else no V
  function vq_arg(v, vt) result(p)
    complex(kind=8), intent(in) :: v(ndim2,ndim2)
    complex(kind=8), intent(inout), target :: vt(ndim2,ndim2)
    type(c_ptr) :: p
    p = c_null_ptr
    if (do_vq) then
      vt = v
      p = c_loc(vt)
    endif
  end function vq_arg

  subroutine calc_eom_static(kq_ind_eom,iq,v,sigma_dmft,sigma_hf,nonlocal)
    integer,intent(in)             :: kq_ind_eom(nkp_eom,nqp)
    integer,intent(in)             :: iq
    complex(kind=8),intent(in)     :: v(ndim2,ndim2)
    complex(kind=8),intent(inout)  :: sigma_dmft(ndim,ndim,-iwfmax_small:iwfmax_small-1)
    complex(kind=8),intent(inout)  :: sigma_hf(ndim,ndim,nkp_eom)
    logical,intent(in)             :: nonlocal
    complex(kind=8), target :: vt(ndim2,ndim2)
    integer(c_int) :: rc, nl
    nl = 0
    if (nonlocal) nl = 1
    rc = c_adga_eom_static(dropin_ctx(.true.), int(iq, c_int), vq_arg(v, vt), nl, sigma_dmft, sigma_hf)
    call dropin_check(rc, 'calc_eom_static')
  end subroutine calc_eom_static

  subroutine calc_eom_dmft(gammawd,gammawm,iwb,sigma_dmft)
    complex(kind=8),intent(in)     :: gammawd(ndim2,maxdim)
    complex(kind=8),intent(in)     :: gammawm(ndim2,maxdim)
    integer,intent(in)             :: iwb
    complex(kind=8),intent(inout)  :: sigma_dmft(ndim,ndim,-iwfmax_small:iwfmax_small-1)
    integer(c_int) :: rc
    if (debug .and. (index(dbgstr,"Dmft_phbar") .ne. 0)) &
      call mpi_stop('adga: the GPU calc_eom_dmft has no Dmft_phbar debug branch', 1)
    rc = c_adga_eom_dmft(dropin_ctx(.true.), int(iwb, c_int), gammawd, sigma_dmft)
    call dropin_check(rc, 'calc_eom_dmft')
  end subroutine calc_eom_dmft

  subroutine calc_eom_dynamic(etaqd,etaqm,gammawd,gammaqd,kq_ind_eom,iwb,iq,v,sigma_nl)
    complex(kind=8),intent(in)     :: etaqd(ndim2,maxdim)
    complex(kind=8),intent(in)     :: etaqm(ndim2,maxdim)
    complex(kind=8),intent(in)     :: gammawd(ndim2,maxdim)
    complex(kind=8),intent(in)     :: gammaqd(ndim2,maxdim)
    integer,intent(in)             :: kq_ind_eom(nkp_eom,nqp)
    integer,intent(in)             :: iwb
    integer,intent(in)             :: iq
    complex(kind=8),intent(in)     :: v(ndim2,ndim2)
    complex(kind=8),intent(inout)  :: sigma_nl(ndim,ndim,-iwfmax_small:iwfmax_small-1,nkp_eom)
    complex(kind=8), target :: vt(ndim2,ndim2)
    integer(c_int) :: rc
    rc = c_adga_eom_dynamic(dropin_ctx(.true.), int(iwb, c_int), int(iq, c_int), etaqd, etaqm, gammawd, gammaqd, &
                            vq_arg(v, vt), sigma_nl)
    call dropin_check(rc, 'calc_eom_dynamic')
  end subroutine calc_eom_dynamic

  subroutine add_siw_dmft(sigma_sum)
    complex(kind=8) :: sigma_sum(ndim, ndim, -iwfmax_small:iwfmax_small-1, nkp_eom)
    integer(c_int) :: rc
    rc = c_adga_add_siw_dmft(dropin_ctx(.true.), sigma_sum)
    call dropin_check(rc, 'add_siw_dmft')
  end subroutine add_siw_dmft

  ! the (re, im) pairs of a(i,i) [i = 1, ndim] (full = .false.) or a(i,j) [j = i, ndim] [i = 1, ndim]
  subroutine pairs(a, full, v, n)
    complex(kind=8), intent(in) :: a(ndim,ndim)
    logical, intent(in) :: full
    real(kind=8), intent(out) :: v(:)
    integer, intent(out) :: n
    integer :: p, q
    n = 0
    do p = 1, ndim
      do q = p, ndim
        if (.not. full .and. q /= p) cycle
        v(n+1) = real(a(p,q))
        v(n+2) = aimag(a(p,q))
        n = n + 2
      enddo
    enddo
  end subroutine pairs

  ! one text file: a header line, then per frequency iw in [lo, hi] the frequency and rows(1:nv, iw)
  subroutine write_rows(name, header, lo, hi, rows, nv)
    character(len=*), intent(in) :: name, header
    integer, intent(in) :: lo, hi, nv
    real(kind=8), intent(in) :: rows(:, lo:)
    integer :: iw
    open(44, file=trim(output_dir)//name, status='unknown')
    write(44,'(A)') header
    do iw = lo, hi
      write(44,'(100F12.6)') iw_data(iw), rows(1:nv, iw)
    enddo
    close(44)
  end subroutine write_rows

  ! rows of a frequency series of n x n blocks: which = 1 sigma-like (n, n, iw), 2 gloc-like (iw, n, n),
  ! 3 the orbital-diagonal siw / giw (iw, n)
  subroutine series_file(name, header, lo, hi, full, which, b3, g3, d2)
    character(len=*), intent(in) :: name, header
    integer, intent(in) :: lo, hi, which
    logical, intent(in) :: full
    complex(kind=8), intent(in), optional :: b3(:,:,lo:), g3(lo:,:,:), d2(lo:,:)
    real(kind=8), allocatable :: rows(:,:)
    complex(kind=8) :: a(ndim,ndim)
    integer :: iw, nv, p
    allocate(rows(2*ndim*ndim, lo:hi))
    nv = 0
    do iw = lo, hi
      a = 0
      select case (which)
      case (1)
        a = b3(:,:,iw)
      case (2)
        a = g3(iw,:,:)
      case default
        do p = 1, ndim
          a(p,p) = d2(iw,p)
        enddo
      end select
      call pairs(a, full, rows(:,iw), nv)
    enddo
    call write_rows(name, header, lo, hi, rows, nv)
    deallocate(rows)
  end subroutine series_file

  subroutine output_eom(sigma_sum, sigma_sum_dmft, sigma_sum_hf, sigma_loc, gloc, nonlocal)
    complex(kind=8), intent(in) :: sigma_sum(ndim, ndim, -iwfmax_small:iwfmax_small-1, nkp_eom)
    complex(kind=8), intent(in) :: sigma_sum_dmft(ndim, ndim, -iwfmax_small:iwfmax_small-1)
    complex(kind=8), intent(in) :: sigma_sum_hf(ndim,ndim,nkp_eom)
    complex(kind=8), intent(in) :: sigma_loc(ndim, ndim, -iwfmax_small:iwfmax_small-1)
    complex(kind=8), intent(in) :: gloc(-iwmax:iwmax-1,ndim,ndim)
    logical,intent(in) :: nonlocal
    character(len=200) :: fname
    real(kind=8) :: vals(3*ndim*ndim)
    integer :: ik, iwf, i, j, nv, lo, hi
    logical :: grid

    lo = -iwfmax_small
    hi = iwfmax_small-1
    grid = nonlocal .and. .not. k_path_eom
    if (grid) call series_file("siw_loc_diag.dat", &
      '# wf, (real(siw_dga_loc(wf,i,i)), imag(siw_dga_loc(wf,i,i))) [i=1,ndim]', lo, hi, .false., 1, b3=sigma_loc)
    call series_file("siw_dmft_rep_diag.dat", &
      '# wf, (real(siw_dmft_rep(wf,i,i)), imag(siw_dmft_rep(wf,i,i))) [i=1,ndim]', lo, hi, .false., 1, b3=sigma_sum_dmft)
    if (grid) call series_file("g_loc_diag.dat", &
      '# wf, (real(giw_dga_loc(iwf,i,i)), imag(giw_dga_loc(iwf,i,i))) [i=1,ndim]', -iwmax, iwmax-1, .false., 2, g3=gloc)
    if (ndim .ge. 2) then
      if (grid) call series_file("siw_loc_full.dat", &
        '# wf, (real(siw_dga_loc(wf,i,j)), imag(siw_dga_loc(iwf,i,j))) [j=i,ndim] [i=1,ndim]', lo, hi, .true., 1, &
        b3=sigma_loc)
      call series_file("siw_dmft_rep_full.dat", &
        '# wf, (real(siw_dmft_rep(wf,i,j)), imag(siw_dmft_rep(iwf,i,j))) [j=i,ndim] [i=1,ndim]', lo, hi, .true., 1, &
        b3=sigma_sum_dmft)
      if (grid) call series_file("g_loc_full.dat", &
        '# wf, (real(giw_dga_loc(wf,i,j)), imag(giw_dga_loc(wf,i,j))) [j=i,ndim] [i=1,ndim]', -iwmax, iwmax-1, &
        .true., 2, g3=gloc)
    endif
    if (.not. nonlocal .or. (verbose .and. (index(verbstr,"Dmft") .ne. 0))) then
      call series_file("siw_dmft_orig.dat", '# wf, (real(siw_dmft(wf,i)), imag(siw_dmft(wf,i))) [i=1,ndim]', &
        lo, hi, .false., 3, d2=siw)
      call series_file("g_dmft_orig.dat", '# wf, (real(giw_dmft(wf,i)), imag(giw_dmft(wf,i))) [i=1,ndim]', &
        -iwmax, iwmax-1, .false., 3, d2=giw)
    endif
    if (.not. nonlocal) return
    if (verbose .and. (index(verbstr,"Siwk") .ne. 0)) then
      do ik = 1, nkp_eom
        write(fname,'(A,F5.3,A,F5.3,A,F5.3,A)') 'siwk_', k_data(1,k_data_eom(ik)), '_', &
          k_data(2,k_data_eom(ik)), '_', k_data(3,k_data_eom(ik)), '.dat'
        call series_file(trim(fname), &
          '# wf, (real(siwk_dga(wf,ik,i,j)), imag(siwk_dga(wf,ik,i,j))) [j=i,ndim] [i=1,ndim]', lo, hi, .true., 1, &
          b3=sigma_sum(:,:,:,ik))
      enddo
    endif
    open(45, file=trim(output_dir)//"siw_all_k.dat", status='unknown')
    write(45,'(A,A)') '# ik, kx, ky, kz, wf, ', &
      '(real(siwk_dga(wf,ik,i,j)), imag(siwk_dga(wf,ik,i,j)), real(siwk_hf(ik,i,j))) [j=i,ndim] [i=1,ndim]'
    do ik = 1, nkp_eom
      do iwf = lo, hi
        nv = 0
        do i = 1, ndim
          do j = i, ndim
            vals(nv+1:nv+3) = [real(sigma_sum(i,j,iwf,ik)), aimag(sigma_sum(i,j,iwf,ik)), real(sigma_sum_hf(i,j,ik))]
            nv = nv + 3
          enddo
        enddo
        write(45,'(I8, 100F12.6)') ik, k_data(:,k_data_eom(ik)), iw_data(iwf), vals(1:nv)
      enddo
    enddo
    close(45)
  end subroutine output_eom

end module eom_module
