! Fortran host check of the GPU hot path through the ISO_C_BINDING layer:
! the (omega, q) loop of src/main.F90:486-799 driven through module adga_gpu,
! and the drop-in lapack_module's inverse_matrix.
!
! Usage: adga_fortran_check <input.bin> <output.bin> <precision: 0 fp64, 1 refined>
! The input/output files have the layout oracle/ref_runner.py writes and reads.
program adga_fortran_check
  use adga_gpu
  use lapack_module
  implicit none
  character(len=512) :: fin, fout, arg
  integer :: n, nkx, nky, nkz, iwmax, iwf, iwbm, nqp, nkp_eom, npts, flags, prec, ierr
  integer :: n2, maxdim, nwu, a, p, e, iwb, slot, i
  real(8) :: beta, mu, err
  complex(8), allocatable :: hk(:,:,:), siw(:,:), giw(:,:), dc(:,:), u(:,:), ut(:,:), ndmft(:)
  complex(8), allocatable :: chil_d(:,:,:), chil_m(:,:,:)
  complex(8), allocatable :: chi_d(:,:,:), chi_m(:,:,:), bub_nl(:,:,:), sig_nl(:,:,:,:), sig_dmft(:,:,:), sig_hf(:,:,:)
  complex(8), allocatable :: x_loc_d(:,:,:), x_loc_m(:,:,:), bub_loc(:,:,:), xld(:,:), xlm(:,:), bl(:,:), zeros(:,:)
  complex(8), allocatable :: am(:,:), bm(:,:)
  integer, allocatable :: q_data(:), k_eom(:), pts_w(:), pts_q(:), wu(:), st(:)
  character(len=200) :: erstr

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg)
  read(arg, *) prec
  open(11, file=trim(fin), access='stream', form='unformatted', status='old')
  read(11) n, nkx, nky, nkz, iwmax, iwf, iwbm, nqp, nkp_eom, npts, flags
  read(11) beta, mu
  n2 = n*n
  maxdim = n2*2*iwf
  allocate(hk(n,n,nkx*nky*nkz), siw(-iwmax:iwmax-1,n), giw(-iwmax:iwmax-1,n), dc(2,n), u(n2,n2), ut(n2,n2), ndmft(n))
  allocate(q_data(nqp), k_eom(nkp_eom), pts_w(npts), pts_q(npts))
  read(11) hk
  read(11) siw
  read(11) giw
  read(11) dc
  read(11) u
  read(11) ut
  read(11) ndmft
  read(11) q_data
  read(11) k_eom
  read(11) pts_w
  read(11) pts_q
  read(11) nwu
  allocate(wu(nwu), chil_d(maxdim,maxdim,nwu), chil_m(maxdim,maxdim,nwu))
  read(11) wu
  do a = 1, nwu
     read(11) chil_d(:,:,a)
     read(11) chil_m(:,:,a)
  enddo
  close(11)

  call adga_gpu_init(n, nkx, nky, nkz, iwmax, iwf, iwbm, beta, mu, hk, siw, giw, dc, u, ut, ndmft, q_data, k_eom, ierr)
  if (ierr /= 0) stop 2
  allocate(chi_d(n2,n2,npts), chi_m(n2,n2,npts), bub_nl(n2,n2,npts), st(npts))
  allocate(sig_nl(n,n,-iwf:iwf-1,nkp_eom), sig_dmft(n,n,-iwf:iwf-1), sig_hf(n,n,nkp_eom))
  allocate(x_loc_d(n2,n2,-iwbm:iwbm), x_loc_m(n2,n2,-iwbm:iwbm), bub_loc(n2,n2,-iwbm:iwbm))
  allocate(xld(n2,n2), xlm(n2,n2), bl(n2,n2), zeros(n2,maxdim))
  x_loc_d = 0; x_loc_m = 0; bub_loc = 0; zeros = 0

  p = 1
  do while (p <= npts)                                   ! one batch per run of equal omega
     iwb = pts_w(p)
     e = p
     do while (e < npts)
        if (pts_w(e+1) /= iwb) exit
        e = e + 1
     enddo
     slot = findloc(wu, iwb, 1)
     call adga_gpu_local(iwb, chil_d(:,:,slot), chil_m(:,:,slot), xld, xlm, bl, ierr)
     if (ierr /= 0) stop 3
     if (pts_q(p) == 1) then
        x_loc_d(:,:,iwb) = x_loc_d(:,:,iwb) + xld
        x_loc_m(:,:,iwb) = x_loc_m(:,:,iwb) + xlm
        bub_loc(:,:,iwb) = bub_loc(:,:,iwb) + bl
     endif
     call adga_gpu_points(pts_q(p:e), prec, chi_d(:,:,p:e), chi_m(:,:,p:e), bub_nl(:,:,p:e), st(p:e), ierr)
     if (ierr /= 0) stop 4
     p = e + 1
  enddo
  call adga_gpu_sigma(sig_nl, sig_dmft, sig_hf, ierr)
  if (ierr /= 0) stop 5
  call adga_gpu_free()

  open(12, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(12) chi_d
  write(12) chi_m
  write(12) bub_nl
  write(12) sig_nl
  write(12) sig_dmft
  write(12) x_loc_d
  write(12) x_loc_m
  write(12) bub_loc
  write(12) zeros
  write(12) zeros
  write(12) zeros
  write(12) sig_hf
  close(12)

  ! drop-in lapack_module: A A^-1 = 1 for a diagonally dominant complex matrix, and the
  ! reference's error contract for a singular one
  allocate(am(37,37), bm(37,37))
  do i = 1, 37
     do a = 1, 37
        am(a,i) = cmplx(sin(0.3d0*a + 0.7d0*i), cos(0.11d0*a*i), kind=8)
     enddo
     am(i,i) = am(i,i) + 40d0
  enddo
  bm = am
  call inverse_matrix(bm, erstr, ierr)
  if (ierr /= 0) stop 6
  bm = matmul(am, bm)
  err = 0
  do i = 1, 37
     bm(i,i) = bm(i,i) - 1d0
  enddo
  err = maxval(abs(bm))
  bm = (1d0, 0d0)
  call inverse_matrix(bm, erstr, ierr)
  write(*,'(a,es10.3,a,i0,2a)') 'inverse_matrix: max|A A^-1 - 1| = ', err, ', singular ierr = ', ierr, &
       ', erstr = ', trim(erstr)
  if (err > 1d-12 .or. ierr == 0 .or. trim(erstr) /= 'ERROR in ZGETRF') stop 7
end program adga_fortran_check
