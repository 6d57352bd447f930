! Test-infrastructure driver that runs the reference's own one-particle routines
! (src/one_particle_quant_module.f90: get_giw 40-79, get_sigma_g_loc 124-180,
! get_nfock 337-361, get_ndga 363-393) on inputs from a binary file.  Compiled by
! oracle/Makefile against the reference sources where they lie, like
! ref_harness.f90.  The self-energy sigma_sum arrives as the caller's array
! (after add_siw_dmft), as main.F90:871 and 890 pass it.
!
! Usage: adga_op_harness <input.bin> <output.bin>
program adga_op_harness
  use parameters_module
  use one_particle_quant_module
  implicit none
  character(len=512) :: fin, fout
  integer :: iw
  complex(kind=8), allocatable :: sigma_sum(:,:,:,:), sigma_loc(:,:,:), gloc(:,:,:)
  real(kind=8), parameter :: pi = 4d0*atan(1d0)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(11, file=trim(fin), access='stream', form='unformatted', status='old')
  read(11) ndim, nkpx, nkpy, nkpz, iwmax, iwfmax_small, nkp_eom
  read(11) beta, mu
  nkp = nkpx*nkpy*nkpz
  ndim2 = ndim*ndim
  verbose = .false.
  verbstr = ''
  text_output = .false.
  allocate(hk(ndim,ndim,nkp), siw(-iwmax:iwmax-1,ndim), giw(-iwmax:iwmax-1,ndim), dc(2,ndim))
  allocate(k_data_eom(nkp_eom), iw_data(-iwmax:iwmax-1))
  allocate(sigma_sum(ndim,ndim,-iwfmax_small:iwfmax_small-1,nkp_eom))
  read(11) hk
  read(11) siw
  read(11) dc
  read(11) k_data_eom
  read(11) sigma_sum
  close(11)
  do iw = -iwmax, iwmax-1
     iw_data(iw) = pi*(2*iw+1)/beta                    ! config_module.f90:510-512
  enddo

  allocate(n_fock(nkp,ndim,ndim), n_dga(ndim), n_dga_k(nkp_eom,ndim,ndim))
  allocate(sigma_loc(ndim,ndim,-iwfmax_small:iwfmax_small-1), gloc(-iwmax:iwmax-1,ndim,ndim))
  call get_giw()
  call get_nfock()
  sigma_loc = 0
  gloc = 0
  if (nkp_eom .eq. nkp) call get_sigma_g_loc(sigma_sum, sigma_loc, gloc)
  call get_ndga(sigma_sum)

  open(12, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(12) giw
  write(12) n_fock
  write(12) sigma_loc
  write(12) gloc
  write(12) n_dga_k
  write(12) n_dga
  close(12)
end program adga_op_harness
