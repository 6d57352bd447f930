! Test-infrastructure driver around the reference's own hot-path modules.
!
! It is compiled (oracle/Makefile) against the reference sources where they lie
! (src/parameters_module.f90, aux.f90, lapack_module.f90, kq_tools.f90,
! mpi_org.F90, one_particle_quant_module.f90, susc_module.f90, eom_module.f90)
! and linked to LAPACK/BLAS (MKL).  The per-(q,omega) loop body of the
! reference is inline in src/main.F90:486-799, so this program sets the
! parameters_module globals from a binary input file and walks the same steps,
! calling the reference's subroutines for every step that is one
! (accumulate_chi0, get_chi0_loc(_inv), calc_chi_qw, inverse_matrix,
! calc_eom_dynamic, calc_eom_dmft, calc_eom_static) and zgemm where main.F90
! calls zgemm.  The input replaces read_vertex (hdf5_module.f90:629-773): the
! local susceptibilities arrive already in its compound-index layout, and V(q)
! arrives per q point in the compound index read_vq (interaction_module.f90:
! 39-123) produces.  Header flags: bit 0 do_vq, bit 1 k_path_eom (the EOM k list
! is not the whole grid: kq_ind_eom from index_kq_eom, main.F90:366-369), bit 2
! do_eom off (a q path, which the reference runs without the EOM, main.F90:341-347),
! bit 3 the debug option "Bubble" (iwstart/iwstop over the large range as
! config_module.f90:467-477 sets them; chi0w and chi0q kept on the box only, as
! main.F90:518, 656 do).
!
! Usage: adga_ref_harness <input.bin> <output.bin>
program adga_ref_harness
  use parameters_module
  use one_particle_quant_module
  use susc_module
  use eom_module
  use lapack_module
  use kq_tools
  implicit none
  character(len=512) :: fin, fout
  integer :: npts, flags, nw, p, iw, iwb, iq, iwf, d, a, b, c, er, cur
  integer :: qlo, nwu, slot
  integer, allocatable :: pts_w(:), pts_q(:), kq_ind(:,:), kq_ind_eom(:,:), wu(:)
  complex(kind=8), allocatable :: chil_d(:,:,:), chil_m(:,:,:)
  complex(kind=8), allocatable :: cwf_d(:,:), cwf_m(:,:), w_inv(:,:,:), chi0w(:,:,:), chi0q(:,:,:)
  complex(kind=8), allocatable :: gam_d(:,:), gam_m(:,:), opg_d(:,:), opg_m(:,:), gamq_d(:,:)
  complex(kind=8), allocatable :: eta_d(:,:), eta_m(:,:), bw_d(:,:), bw_m(:,:), slab(:,:), rect(:,:)
  complex(kind=8), allocatable :: sig_nl(:,:,:,:), sig_dmft(:,:,:), sig_hf(:,:,:), vq(:,:)
  complex(kind=8), allocatable :: x_d(:,:,:), x_m(:,:,:), bub_nl(:,:,:), x_loc_d(:,:,:), x_loc_m(:,:,:)
  complex(kind=8), allocatable :: bub_loc(:,:,:), first_gq(:,:), first_ed(:,:), first_em(:,:), kk(:,:)
  complex(kind=8), allocatable :: chi0nl(:,:,:), vqall(:,:,:)
  complex(kind=8) :: sw(64,64), loc_tmp(64,64), one, zero
  real(kind=8), parameter :: pi = 4d0*atan(1d0)
  character(len=200) :: erstr
  logical :: nonlocal, kpath

  one = (1d0, 0d0)
  zero = (0d0, 0d0)
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(11, file=trim(fin), access='stream', form='unformatted', status='old')
  read(11) ndim, nkpx, nkpy, nkpz, iwmax, iwfmax_small, iwbmax_small, nqp, nkp_eom, npts, flags
  read(11) beta, mu
  nkp = nkpx*nkpy*nkpz
  ndim2 = ndim*ndim
  maxdim = ndim2*2*iwfmax_small
  nw = 2*iwbmax_small+1
  verbose = .false.
  debug = .false.
  dbgstr = ''
  verbstr = ''
  ounit = 0
  do_vq = iand(flags, 1) /= 0
  kpath = iand(flags, 2) /= 0
  do_chi = .true.
  do_eom = iand(flags, 4) == 0
  if (iand(flags, 8) /= 0) then                        ! config_module.f90:467-477
     iwstart = -iwmax+iwbmax_small
     iwstop = iwmax-iwbmax_small-1
  else
     iwstart = -iwfmax_small
     iwstop = iwfmax_small-1
  endif
  nonlocal = .true.
  text_output = .false.
  allocate(hk(ndim,ndim,nkp), siw(-iwmax:iwmax-1,ndim), giw(-iwmax:iwmax-1,ndim), dc(2,ndim))
  allocate(u(ndim2,ndim2), u_tilde(ndim2,ndim2), n_dmft(ndim), q_data(nqp), k_data_eom(nkp_eom))
  allocate(iw_data(-iwmax:iwmax-1))
  allocate(pts_w(npts), pts_q(npts))
  read(11) hk
  read(11) siw
  read(11) giw
  read(11) dc
  read(11) u
  read(11) u_tilde
  read(11) n_dmft
  read(11) q_data
  read(11) k_data_eom
  if (do_vq) then
     allocate(vqall(ndim2,ndim2,nqp))
     read(11) vqall
  endif
  read(11) pts_w
  read(11) pts_q
  ! local susceptibilities only for the omega values the point list uses
  read(11) nwu
  allocate(wu(nwu), chil_d(maxdim,maxdim,nwu), chil_m(maxdim,maxdim,nwu))
  read(11) wu
  do a = 1, nwu
     read(11) chil_d(:,:,a)
     read(11) chil_m(:,:,a)
  enddo
  close(11)
  do iw = -iwmax, iwmax-1
     iw_data(iw) = pi*(2*iw+1)/beta                    ! config_module.f90:510-512
  enddo

  allocate(kq_ind(nkp,nqp), kq_ind_eom(nkp_eom,nqp))
  call index_kq(kq_ind)                                 ! main.F90:364-371
  if (kpath) then
     call index_kq_eom(kq_ind_eom)
  else
     kq_ind_eom = kq_ind
  endif
  if (do_vq) then                                       ! main.F90:262, 271
     allocate(n_fock(nkp,ndim,ndim))
     call get_nfock()
  endif
  allocate(cwf_d(maxdim,maxdim), cwf_m(maxdim,maxdim), w_inv(ndim2,ndim2,-iwfmax_small:iwfmax_small-1))
  allocate(chi0w(ndim2,ndim2,-iwfmax_small:iwfmax_small-1), chi0q(ndim2,ndim2,-iwfmax_small:iwfmax_small-1))
  allocate(chi0nl(ndim2,ndim2,-iwfmax_small:iwfmax_small-1))
  allocate(gam_d(ndim2,maxdim), gam_m(ndim2,maxdim), opg_d(ndim2,maxdim), opg_m(ndim2,maxdim))
  allocate(gamq_d(ndim2,maxdim), eta_d(ndim2,maxdim), eta_m(ndim2,maxdim), slab(ndim2,maxdim), rect(ndim2,maxdim))
  allocate(sig_nl(ndim,ndim,-iwfmax_small:iwfmax_small-1,nkp_eom), sig_dmft(ndim,ndim,-iwfmax_small:iwfmax_small-1))
  allocate(sig_hf(ndim,ndim,nkp_eom), vq(ndim2,ndim2))
  allocate(x_d(ndim2,ndim2,npts), x_m(ndim2,ndim2,npts), bub_nl(ndim2,ndim2,npts))
  allocate(x_loc_d(ndim2,ndim2,-iwbmax_small:iwbmax_small), x_loc_m(ndim2,ndim2,-iwbmax_small:iwbmax_small))
  allocate(bub_loc(ndim2,ndim2,-iwbmax_small:iwbmax_small))
  allocate(first_gq(ndim2,maxdim), first_ed(ndim2,maxdim), first_em(ndim2,maxdim), kk(ndim2,ndim2))
  sig_nl = zero; sig_dmft = zero; sig_hf = zero; vq = zero
  x_d = zero; x_m = zero; bub_nl = zero; x_loc_d = zero; x_loc_m = zero; bub_loc = zero
  cur = iwbmax_small + 3

  do p = 1, npts
     iwb = pts_w(p)
     iq = pts_q(p)
     if (do_vq) then                                      ! main.F90:497-503 (read_vq)
        vq = vqall(:,:,iq)
     else
        vq = zero
     endif
     if (iwb .ne. cur) then                               ! local update, main.F90:506-627
        cur = iwb
        do iwf = -iwfmax_small, iwfmax_small-1
           call get_chi0_loc_inv(iwf, iwb, w_inv(:,:,iwf))
        enddo
        loc_tmp = zero
        do iwf = iwstart, iwstop
           call get_chi0_loc(iwf, iwb, kk)
           loc_tmp(1:ndim2,1:ndim2) = loc_tmp(1:ndim2,1:ndim2) + kk
           if (-iwfmax_small .le. iwf .and. iwf .le. iwfmax_small-1) chi0w(:,:,iwf) = kk
        enddo
        loc_tmp = loc_tmp/(beta**2)
        if (iq .eq. 1) bub_loc(:,:,iwb) = bub_loc(:,:,iwb) + loc_tmp(1:ndim2,1:ndim2)
        slot = findloc(wu, iwb, 1)
        cwf_d = 0.5d0*(chil_d(:,:,slot) + transpose(chil_d(:,:,slot)))
        cwf_m = 0.5d0*(chil_m(:,:,slot) + transpose(chil_m(:,:,slot)))
        do d = 0, 2*iwfmax_small-1
           do a = 1, ndim2
              c = a + d*ndim2
              cwf_d(:,c) = cwf_d(:,c)*w_inv(a,a,d-iwfmax_small)
              cwf_m(:,c) = cwf_m(:,c)*w_inv(a,a,d-iwfmax_small)
              cwf_d(c,c) = cwf_d(c,c) - 1d0
              cwf_m(c,c) = cwf_m(c,c) - 1d0
           enddo
        enddo
        gam_d = zero
        gam_m = zero
        do a = 1, ndim2
           do d = 0, 2*iwfmax_small-1
              gam_d(a,:) = gam_d(a,:) + cwf_d(a+d*ndim2,:)
              gam_m(a,:) = gam_m(a,:) + cwf_m(a+d*ndim2,:)
           enddo
        enddo
        if (iq .eq. 1) then
           call calc_chi_qw(x_loc_d(:,:,iwb), gam_d, chi0w)
           call calc_chi_qw(x_loc_m(:,:,iwb), gam_m, chi0w)
        endif
        opg_d = gam_d
        opg_m = gam_m
        do a = 1, ndim2
           do d = 0, 2*iwfmax_small-1
              opg_d(a,a+d*ndim2) = opg_d(a,a+d*ndim2) + 1d0
              opg_m(a,a+d*ndim2) = opg_m(a,a+d*ndim2) + 1d0
           enddo
        enddo
     endif

     if (do_eom .and. iwb .eq. 0 .and. (iq .eq. 1 .or. do_vq)) &                                           ! main.F90:635-637
        call calc_eom_static(kq_ind_eom, iq, vq, sig_dmft, sig_hf, nonlocal)
     if (do_eom .and. iq .eq. 1) call calc_eom_dmft(gam_d, gam_m, iwb, sig_dmft)                           ! main.F90:639

     chi0q = zero                                                     ! bubble, main.F90:644-657
     do iwf = iwstart, iwstop
        kk = zero
        do a = 1, nkp
           call accumulate_chi0(a, kq_ind(a,iq), iwf, iwb, kk)
        enddo
        kk = kk/dble(nkp)
        bub_nl(:,:,p) = bub_nl(:,:,p) + kk
        if (-iwfmax_small .le. iwf .and. iwf .le. iwfmax_small-1) chi0q(:,:,iwf) = kk
     enddo
     bub_nl(:,:,p) = bub_nl(:,:,p)/(beta**2) - loc_tmp(1:ndim2,1:ndim2)
     chi0nl = chi0q - chi0w
     call calc_chi_qw(x_d(:,:,p), gam_d, chi0nl)
     call calc_chi_qw(x_m(:,:,p), gam_m, chi0nl)

     allocate(bw_d(maxdim,maxdim), bw_m(maxdim,maxdim))               ! BSE build, main.F90:673-720
     bw_d = zero
     bw_m = zero
     gamq_d = zero
     do d = 0, 2*iwfmax_small-1
        iwf = d - iwfmax_small
        qlo = d*ndim2
        do b = 1, ndim2
           sw(1:ndim2,b) = chi0q(:,b,iwf)*w_inv(b,b,iwf)
           sw(b,b) = sw(b,b) - 1d0
        enddo
        slab = cwf_d(qlo+1:qlo+ndim2,:)
        call zgemm('n','n',ndim2,maxdim,ndim2,one,sw,64,slab,ndim2,zero,rect,ndim2)
        bw_d(qlo+1:qlo+ndim2,:) = -rect
        gamq_d = gamq_d + rect
        slab = cwf_m(qlo+1:qlo+ndim2,:)
        call zgemm('n','n',ndim2,maxdim,ndim2,one,sw,64,slab,ndim2,zero,rect,ndim2)
        bw_m(qlo+1:qlo+ndim2,:) = -rect
        if (do_vq) then                                               ! main.F90:710-719
           sw(1:ndim2,1:ndim2) = 2d0/(beta**2)*matmul(chi0q(:,:,iwf), vq)
           call zgemm('n','n',ndim2,maxdim,ndim2,one,sw,64,opg_d,ndim2,zero,rect,ndim2)
           bw_d(qlo+1:qlo+ndim2,:) = bw_d(qlo+1:qlo+ndim2,:) - rect
        endif
     enddo
     do a = 1, maxdim                                                 ! inversion, main.F90:726-743
        bw_d(a,a) = bw_d(a,a) + 1d0
        bw_m(a,a) = bw_m(a,a) + 1d0
     enddo
     call inverse_matrix(bw_d, erstr, er)
     if (er .ne. 0) stop 3
     call inverse_matrix(bw_m, erstr, er)
     if (er .ne. 0) stop 3
     do a = 1, maxdim
        bw_d(a,a) = bw_d(a,a) - 1d0
        bw_m(a,a) = bw_m(a,a) - 1d0
     enddo
     call zgemm('n','n',ndim2,maxdim,maxdim,one,opg_d,ndim2,bw_d,maxdim,zero,eta_d,ndim2)   ! eta, main.F90:749-757
     call zgemm('n','n',ndim2,maxdim,maxdim,one,opg_m,ndim2,bw_m,maxdim,zero,eta_m,ndim2)
     deallocate(bw_d, bw_m)
     call calc_chi_qw(x_d(:,:,p), eta_d, chi0q)                       ! main.F90:766-769
     call calc_chi_qw(x_m(:,:,p), eta_m, chi0q)
     if (do_eom) call calc_eom_dynamic(eta_d, eta_m, gam_d, gamq_d, kq_ind_eom, iwb, iq, vq, sig_nl)  ! main.F90:774-781
     if (p .eq. 1) then
        first_gq = gamq_d
        first_ed = eta_d
        first_em = eta_m
     endif
  enddo

  open(12, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(12) x_d
  write(12) x_m
  write(12) bub_nl
  write(12) sig_nl
  write(12) sig_dmft
  write(12) x_loc_d
  write(12) x_loc_m
  write(12) bub_loc
  write(12) first_gq
  write(12) first_ed
  write(12) first_em
  write(12) sig_hf
  close(12)
end program adga_ref_harness
