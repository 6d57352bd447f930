! Batched Fortran driver of the GPU hot path: the calls a Fortran host makes in
! place of the per-(q,omega) loop body (src/main.F90:486-799).  Arguments are
! the reference's own arrays with their Fortran shapes (the parameters_module
! globals are passed explicitly), so the module does not depend on the
! reference's modules.
module adga_gpu
  use, intrinsic :: iso_c_binding
  use adga_c
  implicit none
  private
  public :: adga_gpu_init, adga_gpu_local, adga_gpu_points, adga_gpu_sigma, adga_gpu_free, adga_gpu_option, &
            adga_gpu_giw, adga_gpu_nfock, adga_gpu_sigma_g_loc, adga_gpu_ndga, ADGA_PREC_FP64, ADGA_PREC_REFINED

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: n_ = 0

contains

  ! hk(n,n,nkp), siw/giw(-iwmax:iwmax-1,n), dc(2,n), u/u_tilde(n^2,n^2), n_dmft(n), q_data(nqp), k_data_eom(nkp_eom)
  subroutine adga_gpu_init(n, nkpx, nkpy, nkpz, iwmax, iwfmax_small, iwbmax_small, beta, mu, hk, siw, giw, dc, &
                           u, u_tilde, n_dmft, q_data, k_data_eom, ierr)
    integer, intent(in) :: n, nkpx, nkpy, nkpz, iwmax, iwfmax_small, iwbmax_small
    real(8), intent(in) :: beta, mu
    complex(8), intent(in) :: hk(:,:,:), siw(:,:), giw(:,:), dc(:,:), u(:,:), u_tilde(:,:), n_dmft(:)
    integer, intent(in) :: q_data(:), k_data_eom(:)
    integer, intent(out) :: ierr
    type(adga_dims) :: d
    integer(c_int), allocatable :: q(:), k(:)

    d%ndim = n; d%nkpx = nkpx; d%nkpy = nkpy; d%nkpz = nkpz; d%iwmax = iwmax
    d%iwfmax_small = iwfmax_small; d%iwbmax_small = iwbmax_small
    d%nqp = size(q_data); d%nkp_eom = size(k_data_eom); d%beta = beta; d%mu = mu
    allocate(q(size(q_data)), k(size(k_data_eom)))
    q = q_data
    k = k_data_eom
    ierr = int(c_adga_ctx_create(d, q, k, hk, siw, giw, dc, u, u_tilde, n_dmft, ctx))
    n_ = n
  end subroutine adga_gpu_init

  subroutine adga_gpu_option(key, value, ierr)
    character(len=*), intent(in) :: key
    real(8), intent(in) :: value
    integer, intent(out) :: ierr
    ierr = int(c_adga_set_option(ctx, trim(key)//c_null_char, real(value, c_double)))
  end subroutine adga_gpu_option

  ! Local update for omega = iwb: chi_loc_dens/magn(maxdim,maxdim) as read_vertex returns them.
  ! chi_loc_dens_w/magn_w(n^2,n^2) and bubble_loc_w(n^2,n^2) are what the reference accumulates for iq == 1.
  subroutine adga_gpu_local(iwb, chi_loc_dens, chi_loc_magn, chi_loc_dens_w, chi_loc_magn_w, bubble_loc_w, ierr)
    integer, intent(in) :: iwb
    complex(8), intent(in) :: chi_loc_dens(:,:), chi_loc_magn(:,:)
    complex(8), intent(out) :: chi_loc_dens_w(:,:), chi_loc_magn_w(:,:), bubble_loc_w(:,:)
    integer, intent(out) :: ierr
    ierr = int(c_adga_set_local_omega(ctx, int(iwb, c_int), chi_loc_dens, chi_loc_magn, chi_loc_dens_w, &
                                      chi_loc_magn_w, bubble_loc_w))
  end subroutine adga_gpu_local

  ! The per-point work for the q indices iq(:) (1-based into q_data) at the current omega.
  ! chi_dens/chi_magn/bubble_nl(n^2,n^2,size(iq)) receive chi_qw_dens/magn and bubble_nl of the points.
  subroutine adga_gpu_points(iq, precision, chi_dens, chi_magn, bubble_nl, status, ierr)
    integer, intent(in) :: iq(:), precision
    complex(8), intent(out) :: chi_dens(:,:,:), chi_magn(:,:,:), bubble_nl(:,:,:)
    integer, intent(out) :: status(:)
    integer, intent(out) :: ierr
    integer(c_int), allocatable :: q(:), st(:)
    allocate(q(size(iq)), st(size(iq)))
    q = iq
    ierr = int(c_adga_run_points(ctx, int(size(iq), c_int), q, int(precision, c_int), chi_dens, chi_magn, &
                                 bubble_nl, st))
    status = st
  end subroutine adga_gpu_points

  ! sigma_nl(n,n,-iwfmax_small:iwfmax_small-1,nkp_eom), sigma_dmft(n,n,...), sigma_hf(n,n,nkp_eom)
  subroutine adga_gpu_sigma(sigma_nl, sigma_dmft, sigma_hf, ierr)
    complex(8), intent(out) :: sigma_nl(:,:,:,:), sigma_dmft(:,:,:), sigma_hf(:,:,:)
    integer, intent(out) :: ierr
    ierr = int(c_adga_get_sigma(ctx, sigma_nl, sigma_dmft, sigma_hf))
  end subroutine adga_gpu_sigma

  ! the one-particle routines of src/one_particle_quant_module.f90 on the context's hk, siw, dc, mu,
  ! beta and k_data_eom.  giw(-iwmax:iwmax-1,n) (get_giw; also replaces the context's giw)
  subroutine adga_gpu_giw(giw, ierr)
    complex(8), intent(out) :: giw(:,:)
    integer, intent(out) :: ierr
    ierr = int(c_adga_get_giw(ctx, giw))
  end subroutine adga_gpu_giw

  ! n_fock(nkp,n,n) (get_nfock)
  subroutine adga_gpu_nfock(n_fock, ierr)
    complex(8), intent(out) :: n_fock(:,:,:)
    integer, intent(out) :: ierr
    ierr = int(c_adga_get_nfock(ctx, n_fock))
  end subroutine adga_gpu_nfock

  ! sigma_sum(n,n,-iwfmax_small:iwfmax_small-1,nkp), sigma_loc(n,n,-iwfmax_small:iwfmax_small-1),
  ! gloc(-iwmax:iwmax-1,n,n) (get_sigma_g_loc)
  subroutine adga_gpu_sigma_g_loc(sigma_sum, sigma_loc, gloc, ierr)
    complex(8), intent(in) :: sigma_sum(:,:,:,:)
    complex(8), intent(out) :: sigma_loc(:,:,:), gloc(:,:,:)
    integer, intent(out) :: ierr
    ierr = int(c_adga_get_sigma_g_loc(ctx, sigma_sum, sigma_loc, gloc))
  end subroutine adga_gpu_sigma_g_loc

  ! sigma_sum(n,n,-iwfmax_small:iwfmax_small-1,nkp_eom), n_dga_k(nkp_eom,n,n), n_dga(n) (get_ndga)
  subroutine adga_gpu_ndga(sigma_sum, n_dga_k, n_dga, ierr)
    complex(8), intent(in) :: sigma_sum(:,:,:,:)
    complex(8), intent(out) :: n_dga_k(:,:,:), n_dga(:)
    integer, intent(out) :: ierr
    ierr = int(c_adga_get_ndga(ctx, sigma_sum, n_dga_k, n_dga))
  end subroutine adga_gpu_ndga

  subroutine adga_gpu_free()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = c_adga_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine adga_gpu_free

end module adga_gpu
