! ISO_C_BINDING interfaces to libadga_amd.so (include/adga_amd.h).
!
! Complex arrays cross the boundary as complex(c_double_complex), which has the
! layout of the reference's complex(kind=8): interleaved (re, im), column-major.
module adga_c
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: adga_dims, c_adga_ctx_create, c_adga_set_option, c_adga_set_local_omega, c_adga_run_points, &
            c_adga_get_sigma, c_adga_reset_sigma, c_adga_get_stats, c_adga_last_error, c_adga_ctx_destroy, &
            c_adga_inverse_matrix_z, c_adga_get_giw, c_adga_get_nfock, c_adga_get_sigma_g_loc, c_adga_get_ndga, &
            c_adga_set_vq, c_adga_bubble_q, c_adga_chi0_loc, c_adga_chi_qw, c_adga_eom_dynamic, c_adga_eom_dmft, &
            c_adga_eom_static, c_adga_get_gkiw, c_adga_get_ndmft, c_adga_add_siw_dmft, &
            c_adga_set_local_omega_box, c_adga_inverse_matrix_z_noalloc, c_adga_set_tangents, &
            c_adga_set_local_omega_tangent, c_adga_run_points_tangent, c_adga_get_sigma_tangent, &
            c_adga_lambda_create, c_adga_lambda_sums, c_adga_lambda_target, c_adga_lambda_solve, c_adga_lambda_chi, &
            c_adga_lambda_last_error, c_adga_lambda_destroy, &
            ADGA_OK, ADGA_E_SINGULAR, ADGA_PREC_FP64, ADGA_PREC_REFINED

  integer(c_int), parameter :: ADGA_OK = 0, ADGA_E_SINGULAR = 3
  integer(c_int), parameter :: ADGA_PREC_FP64 = 0, ADGA_PREC_REFINED = 1

  type, bind(c) :: adga_dims
     integer(c_int) :: ndim, nkpx, nkpy, nkpz, iwmax, iwfmax_small, iwbmax_small, nqp, nkp_eom
     real(c_double) :: beta, mu
  end type adga_dims

  interface
     integer(c_int) function c_adga_ctx_create(dims, q_data, k_data_eom, hk, siw, giw, dc, u, u_tilde, n_dmft, ctx) &
          bind(c, name='adga_ctx_create')
       import :: c_int, c_ptr, c_double_complex, adga_dims
       type(adga_dims), intent(in) :: dims
       integer(c_int), intent(in) :: q_data(*), k_data_eom(*)
       complex(c_double_complex), intent(in) :: hk(*), siw(*), giw(*), dc(*), u(*), u_tilde(*), n_dmft(*)
       type(c_ptr), intent(out) :: ctx
     end function c_adga_ctx_create

     integer(c_int) function c_adga_set_option(ctx, key, value) bind(c, name='adga_set_option')
       import :: c_int, c_ptr, c_char, c_double
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: key(*)
       real(c_double), value :: value
     end function c_adga_set_option

     integer(c_int) function c_adga_set_vq(ctx, v_q) bind(c, name='adga_set_vq')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: v_q(*)
     end function c_adga_set_vq

     integer(c_int) function c_adga_set_local_omega(ctx, iwb, chi_loc_dens, chi_loc_magn, chi_loc_dens_out, &
          chi_loc_magn_out, bubble_loc_out) bind(c, name='adga_set_local_omega')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: iwb
       complex(c_double_complex), intent(in) :: chi_loc_dens(*), chi_loc_magn(*)
       complex(c_double_complex), intent(out) :: chi_loc_dens_out(*), chi_loc_magn_out(*), bubble_loc_out(*)
     end function c_adga_set_local_omega

     integer(c_int) function c_adga_run_points(ctx, npts, iq, precision, chi_dens_out, chi_magn_out, bubble_nl_out, &
          status) bind(c, name='adga_run_points')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: npts, precision
       integer(c_int), intent(in) :: iq(*)
       complex(c_double_complex), intent(out) :: chi_dens_out(*), chi_magn_out(*), bubble_nl_out(*)
       integer(c_int), intent(out) :: status(*)
     end function c_adga_run_points

     integer(c_int) function c_adga_get_sigma(ctx, sigma_nl, sigma_dmft, sigma_hf) bind(c, name='adga_get_sigma')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(out) :: sigma_nl(*), sigma_dmft(*), sigma_hf(*)
     end function c_adga_get_sigma

     integer(c_int) function c_adga_reset_sigma(ctx) bind(c, name='adga_reset_sigma')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function c_adga_reset_sigma

     integer(c_int) function c_adga_get_stats(ctx, out, nout) bind(c, name='adga_get_stats')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(*)
       integer(c_int), value :: nout
     end function c_adga_get_stats

     integer(c_int) function c_adga_last_error(ctx, buf, len) bind(c, name='adga_last_error')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: len
     end function c_adga_last_error

     integer(c_int) function c_adga_ctx_destroy(ctx) bind(c, name='adga_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function c_adga_ctx_destroy

     integer(c_int) function c_adga_inverse_matrix_z(n, m, ierr) bind(c, name='adga_inverse_matrix_z')
       import :: c_int, c_double_complex
       integer(c_int), value :: n
       complex(c_double_complex), intent(inout) :: m(*)
       integer(c_int), intent(out) :: ierr
     end function c_adga_inverse_matrix_z

     integer(c_int) function c_adga_get_giw(ctx, giw) bind(c, name='adga_get_giw')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(out) :: giw(*)
     end function c_adga_get_giw

     integer(c_int) function c_adga_get_nfock(ctx, n_fock) bind(c, name='adga_get_nfock')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(out) :: n_fock(*)
     end function c_adga_get_nfock

     integer(c_int) function c_adga_get_sigma_g_loc(ctx, sigma_sum, sigma_loc, gloc) bind(c, name='adga_get_sigma_g_loc')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: sigma_sum(*)
       complex(c_double_complex), intent(out) :: sigma_loc(*), gloc(*)
     end function c_adga_get_sigma_g_loc

     integer(c_int) function c_adga_get_ndga(ctx, sigma_sum, n_dga_k, n_dga) bind(c, name='adga_get_ndga')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: sigma_sum(*)
       complex(c_double_complex), intent(out) :: n_dga_k(*), n_dga(*)
     end function c_adga_get_ndga

     ! module-level entry points (one call per call of a reference module procedure)
     integer(c_int) function c_adga_bubble_q(ctx, iwb, q_k, chi0q) bind(c, name='adga_bubble_q')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: iwb, q_k
       complex(c_double_complex), intent(out) :: chi0q(*)
     end function c_adga_bubble_q

     integer(c_int) function c_adga_chi0_loc(ctx, iwb, chi0w_diag, chi0w_inv_diag) bind(c, name='adga_chi0_loc')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: iwb
       complex(c_double_complex), intent(out) :: chi0w_diag(*), chi0w_inv_diag(*)
     end function c_adga_chi0_loc

     integer(c_int) function c_adga_chi_qw(ctx, interm3, chi0_sum, chi_qw) bind(c, name='adga_chi_qw')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: interm3(*), chi0_sum(*)
       complex(c_double_complex), intent(inout) :: chi_qw(*)
     end function c_adga_chi_qw

     integer(c_int) function c_adga_eom_dynamic(ctx, iwb, iq, etaqd, etaqm, gammawd, gammaqd, v, sigma_nl) &
          bind(c, name='adga_eom_dynamic')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx, v
       integer(c_int), value :: iwb, iq
       complex(c_double_complex), intent(in) :: etaqd(*), etaqm(*), gammawd(*), gammaqd(*)
       complex(c_double_complex), intent(inout) :: sigma_nl(*)
     end function c_adga_eom_dynamic

     integer(c_int) function c_adga_eom_dmft(ctx, iwb, gammawd, sigma_dmft) bind(c, name='adga_eom_dmft')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: iwb
       complex(c_double_complex), intent(in) :: gammawd(*)
       complex(c_double_complex), intent(inout) :: sigma_dmft(*)
     end function c_adga_eom_dmft

     integer(c_int) function c_adga_eom_static(ctx, iq, v, nonlocal, sigma_dmft, sigma_hf) bind(c, name='adga_eom_static')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx, v
       integer(c_int), value :: iq, nonlocal
       complex(c_double_complex), intent(inout) :: sigma_dmft(*), sigma_hf(*)
     end function c_adga_eom_static

     integer(c_int) function c_adga_get_gkiw(ctx, ik, iw, skiw, gkiw) bind(c, name='adga_get_gkiw')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx, skiw
       integer(c_int), value :: ik, iw
       complex(c_double_complex), intent(out) :: gkiw(*)
     end function c_adga_get_gkiw

     integer(c_int) function c_adga_get_ndmft(ctx, n_dmft) bind(c, name='adga_get_ndmft')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(out) :: n_dmft(*)
     end function c_adga_get_ndmft

     integer(c_int) function c_adga_add_siw_dmft(ctx, sigma_sum) bind(c, name='adga_add_siw_dmft')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(inout) :: sigma_sum(*)
     end function c_adga_add_siw_dmft

     integer(c_int) function c_adga_set_local_omega_box(ctx, iwb, iwf_box, chi_loc_dens, chi_loc_magn, &
          chi_loc_dens_out, chi_loc_magn_out, bubble_loc_out) bind(c, name='adga_set_local_omega_box')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: iwb, iwf_box
       complex(c_double_complex), intent(in) :: chi_loc_dens(*), chi_loc_magn(*)
       complex(c_double_complex), intent(out) :: chi_loc_dens_out(*), chi_loc_magn_out(*), bubble_loc_out(*)
     end function c_adga_set_local_omega_box

     integer(c_int) function c_adga_inverse_matrix_z_noalloc(n, m, ierr) bind(c, name='adga_inverse_matrix_z_noalloc')
       import :: c_int, c_double_complex
       integer(c_int), value :: n
       complex(c_double_complex), intent(inout) :: m(*)
       integer(c_int), intent(out) :: ierr
     end function c_adga_inverse_matrix_z_noalloc

     ! derivatives along ndir directions of the inputs (no counterpart in the reference)
     integer(c_int) function c_adga_set_tangents(ctx, ndir, dhk, dsiw, dgiw, dmu, du, du_tilde) &
          bind(c, name='adga_set_tangents')
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndir
       complex(c_double_complex), intent(in) :: dhk(*), dsiw(*), dgiw(*), du(*), du_tilde(*)
       real(c_double), intent(in) :: dmu(*)
     end function c_adga_set_tangents

     integer(c_int) function c_adga_set_local_omega_tangent(ctx, dchi_loc_dens, dchi_loc_magn, dchi_loc_dens_out, &
          dchi_loc_magn_out, dbubble_loc_out) bind(c, name='adga_set_local_omega_tangent')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: dchi_loc_dens(*), dchi_loc_magn(*)
       complex(c_double_complex), intent(out) :: dchi_loc_dens_out(*), dchi_loc_magn_out(*), dbubble_loc_out(*)
     end function c_adga_set_local_omega_tangent

     integer(c_int) function c_adga_run_points_tangent(ctx, npts, iq, chi_dens_out, chi_magn_out, bubble_nl_out, &
          dchi_dens_out, dchi_magn_out, dbubble_nl_out, status) bind(c, name='adga_run_points_tangent')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: npts
       integer(c_int), intent(in) :: iq(*)
       complex(c_double_complex), intent(out) :: chi_dens_out(*), chi_magn_out(*), bubble_nl_out(*)
       complex(c_double_complex), intent(out) :: dchi_dens_out(*), dchi_magn_out(*), dbubble_nl_out(*)
       integer(c_int), intent(out) :: status(*)
     end function c_adga_run_points_tangent

     integer(c_int) function c_adga_get_sigma_tangent(ctx, dsigma_nl) bind(c, name='adga_get_sigma_tangent')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(out) :: dsigma_nl(*)
     end function c_adga_get_sigma_tangent

     ! lambda correction of the final susceptibility (no counterpart in the reference)
     integer(c_int) function c_adga_lambda_create(ndim, nqp, iwbmax, beta, chi_qw, chi_loc, h) &
          bind(c, name='adga_lambda_create')
       import :: c_int, c_ptr, c_double, c_double_complex
       integer(c_int), value :: ndim, nqp, iwbmax
       real(c_double), value :: beta
       complex(c_double_complex), intent(in) :: chi_qw(*), chi_loc(*)
       type(c_ptr), intent(out) :: h
     end function c_adga_lambda_create

     integer(c_int) function c_adga_lambda_sums(h, nl, lambdas, s, ds) bind(c, name='adga_lambda_sums')
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: nl
       real(c_double), intent(in) :: lambdas(*)
       complex(c_double_complex), intent(out) :: s(*), ds(*)
     end function c_adga_lambda_sums

     integer(c_int) function c_adga_lambda_target(h, target) bind(c, name='adga_lambda_target')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: target
     end function c_adga_lambda_target

     integer(c_int) function c_adga_lambda_solve(h, target, tol, maxit, lambda, iters) bind(c, name='adga_lambda_solve')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: target, tol
       integer(c_int), value :: maxit
       real(c_double), intent(out) :: lambda
       integer(c_int), intent(out) :: iters
     end function c_adga_lambda_solve

     integer(c_int) function c_adga_lambda_chi(h, lambda, chi_phys) bind(c, name='adga_lambda_chi')
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: h
       real(c_double), value :: lambda
       complex(c_double_complex), intent(out) :: chi_phys(*)
     end function c_adga_lambda_chi

     integer(c_int) function c_adga_lambda_last_error(h, buf, len) bind(c, name='adga_lambda_last_error')
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: h
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: len
     end function c_adga_lambda_last_error

     integer(c_int) function c_adga_lambda_destroy(h) bind(c, name='adga_lambda_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function c_adga_lambda_destroy
  end interface
end module adga_c
