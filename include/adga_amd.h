/*
 * adga_amd.h - C ABI of the MI355X (gfx950) implementation of ADGA's per-(q,omega) hot path:
 * the chi0 bubble, the density/magnetic Bethe-Salpeter solves, the chi(q,omega) contraction and
 * the equation of motion for Sigma_DGA(k,nu).
 *
 * Conventions (all entry points):
 *   - complex arrays are interleaved (re, im) doubles, i.e. Fortran complex(kind=8), passed as
 *     `double *` and laid out column-major with the reference's Fortran shapes (first index
 *     fastest); frequency axes keep the reference's offsets (siw(-iwmax:iwmax-1, ndim) is passed
 *     as a (2*iwmax, ndim) array whose row iw+iwmax holds nu_iw);
 *   - integers are int32; q and k indices are the reference's 1-based k-grid indices
 *     (ik = 1 + kz + ky*nkz + kx*nky*nkz, src/kq_tools.f90:119-125);
 *   - every entry point returns an ADGA_* status code and never throws across the boundary;
 *     adga_last_error() gives the message of the last failure on a context.
 * Arrays are caller-owned host memory; the library copies what it needs to the GPU.
 */
#ifndef ADGA_AMD_H
#define ADGA_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ADGA_OK = 0,
  ADGA_E_ARG = 1,        /* invalid argument or size */
  ADGA_E_HIP = 2,        /* HIP runtime failure (no device, launch failure, ...) */
  ADGA_E_SINGULAR = 3,   /* exactly zero pivot in a BSE LU: the reference's "ERROR in ZGETRF" */
  ADGA_E_NOCONV = 4,     /* refined mode with the option "fallback" off: residual above tolerance */
  ADGA_E_NONFINITE = 5,  /* NaN/Inf met in a BSE matrix or residual */
  ADGA_E_STATE = 6,      /* call out of order (e.g. run_points before set_local_omega) */
  ADGA_E_NOMEM = 7       /* device allocation failed */
};

/* largest n adga_inverse_matrix_z takes (larger: ADGA_E_ARG; the Fortran lapack_module falls back to the
   host LAPACK there) */
#define ADGA_INVERSE_MAX_N 4096

enum {
  ADGA_PREC_FP64 = 0,    /* complex128 LU and solve */
  ADGA_PREC_REFINED = 1  /* complex64 LU, FP64 residuals and iterative refinement */
};

typedef struct adga_ctx adga_ctx;

/* Sizes and scalars of the reference's parameters_module (src/parameters_module.f90:43-73). */
typedef struct adga_dims {
  int ndim;                 /* orbitals n */
  int nkpx, nkpy, nkpz;     /* k grid (nkp = product) */
  int iwmax;                /* 1P frequencies: siw/giw hold 2*iwmax entries */
  int iwfmax_small;         /* fermionic box: Nf = 2*iwfmax_small, BSE size N = n^2*Nf */
  int iwbmax_small;         /* bosonic box: omega in [-iwbmax_small, iwbmax_small] */
  int nqp;                  /* length of the q list */
  int nkp_eom;              /* length of the EOM k list */
  double beta, mu;
} adga_dims;

/*
 * Creates a context on the current HIP device and uploads the one-particle data (the hidden
 * inputs of the hot routines: hk, siw, giw, dc, u, u_tilde, n_dmft, q_data, k_data_eom), then
 * builds the device table G(k, nu) for every nu the loop needs.
 *   hk(ndim,ndim,nkp), siw/giw(2*iwmax,ndim), dc(2,ndim), u/u_tilde(ndim^2,ndim^2), n_dmft(ndim),
 *   q_data(nqp), k_data_eom(nkp_eom).
 * Replaces: the setup the reference does before its loop (src/main.F90:174-389) for the data the
 * hot routines read through parameters_module.
 */
int adga_ctx_create(const adga_dims *dims, const int *q_data, const int *k_data_eom, const double *hk,
                    const double *siw, const double *giw, const double *dc, const double *u,
                    const double *u_tilde, const double *n_dmft, adga_ctx **ctx);

/*
 * Options: "batch" (points per device batch, default 256), "refine_tol" (refined mode: relative
 * residual max|r|/max|b| every BSE solve must reach, default 1e-12), "refine_maxit" (default 12;
 * a matrix that has not converged by then, or whose residual a correction did not halve, or whose
 * complex64 LU met a zero pivot, is solved again in complex128 unless "fallback" is 0; its final
 * status then depends only on the complex128 solve), "do_eom" (1: the default; 0: no
 * equation of motion, as the reference's do_eom = .false. for a q path),
 * "bubble" (0: direct k sum per point; 1: all q of the current omega at once by
 * k-grid DFT, cached per omega; 2: DFT when a batch has >= 64 points or the cache is valid, the
 * default), "eom" (0: direct sum per batch; 1: m_tot accumulated on the q grid and convolved with
 * G by k-grid DFT when omega changes or the self-energy is read, the default), "lu_streams" (1: each
 * batch's LU on one stream; 2 .. 4: the batch split into that many parts on as many streams, each one
 * operation behind the previous, the large Schur updates on the whole batch; 0, the default: chosen per
 * device and batch shape (precision, BSE size, matrices in the batch): the shape's first three
 * factorisations in the process, in any context, run with 1, 2 and 3 streams, each timed, the later
 * ones with the fastest; results are identical for every value),
 * "bubble_range" (0: bubble_loc and bubble_nl sum chi0 over the vertex box, the default; 1: over every
 * nu for which G(nu - omega) exists, [-iwmax + iwbmax_small, iwmax - iwbmax_small), the reference's
 * debug option "Bubble" (src/config_module.f90:467-477); 2: as 1, and the bubble_loc returned by
 * adga_set_local_omega also holds the Matsubara tail beyond that range, -1/beta sum g_i(nu) g_j(nu-omega)
 * with g expanded to third order in 1/(i nu), its moments fitted to the two highest frequencies of giw,
 * in FP64 on the device (no counterpart in the reference; bubble_nl is unchanged by it, its tail being
 * O(nu^-4)); 1 and 2 widen the device G table to all 2*iwmax frequencies and need adga_set_local_omega
 * again before the next points; not with tangent directions).
 * Returns ADGA_E_ARG for an unknown key or value.
 */
int adga_set_option(adga_ctx *ctx, const char *key, double value);

/*
 * Non-local interaction V(q) for every q of the list: v_q(ndim^2, ndim^2, nqp) in the compound index
 * read_vq returns (v(i1 = {k i}, i2 = {j l}) = V_ijkl(q), i.e. i1 = (i-1)*ndim + k, i2 = (l-1)*ndim + j
 * in 1-based Fortran terms).  With V set, adga_run_points adds the density-channel BSE term
 * -2/beta^2 chi0^q(nu) V(q) (1 + gamma^w_d), the V terms of m_tot and, at omega = 0, the non-local Hartree
 * (at the first q of the list) and Fock terms of sigma_hf, as the reference does with do_vq.  NULL
 * removes V.  ADGA_E_STATE while the EOM of an omega is pending (call adga_get_sigma first).
 * Replaces: read_vq per point (src/interaction_module.f90:39-123, src/main.F90:497-503) with do_vq set.
 */
int adga_set_vq(adga_ctx *ctx, const double *v_q);

/*
 * Local update for bosonic frequency iwb: chi0^w and its inverse, chi0^w F_r = chi_loc_r chi0^w^-1 - 1
 * after the transpose symmetrisation, gamma^w_r and 1 + gamma^w_r.
 *   chi_loc_dens/magn(N,N): what read_vertex returns (src/hdf5_module.f90:708-770).
 *   Outputs (may be NULL): chi_loc_dens_out/magn_out(ndim^2,ndim^2) = calc_chi_qw(gamma^w, chi0^w)
 *   and bubble_loc_out(ndim^2,ndim^2) = sum_nu chi0^w / beta^2, which the reference accumulates when
 *   iq == 1.
 * Replaces: src/main.F90:506-625 (get_chi0_loc_inv, get_chi0_loc, symmetrisation, chi0^w F, gamma^w).
 */
int adga_set_local_omega(adga_ctx *ctx, int iwb, const double *chi_loc_dens, const double *chi_loc_magn,
                         double *chi_loc_dens_out, double *chi_loc_magn_out, double *bubble_loc_out);

/*
 * adga_set_local_omega with a fermionic box of this omega's own: nu in [-iwf_box, iwf_box), 1 <= iwf_box <=
 * iwfmax_small, so BSE size N(omega) = ndim^2 * 2 * iwf_box and chi_loc_dens/magn(N(omega), N(omega)).  The
 * following adga_run_points calls solve at that size; the self-energies keep the context's box (the
 * frequencies outside this omega's box receive nothing from it).  Every quantity of the omega is the
 * reference's with iwfmax_small = iwf_box.
 * Replaces: the loop body at one omega with a box that shrinks with |omega| (the reference has one global
 * iwfmax_small, src/config_module.f90:469).
 */
int adga_set_local_omega_box(adga_ctx *ctx, int iwb, int iwf_box, const double *chi_loc_dens,
                             const double *chi_loc_magn, double *chi_loc_dens_out, double *chi_loc_magn_out,
                             double *bubble_loc_out);

/*
 * Runs the per-point work for npts q points at the current bosonic frequency: bubble, BSE build,
 * both channel solves, eta, the chi contraction and the dynamic EOM (Sigma_nl accumulated on the
 * device), plus the local EOM for every point with iq == 1 (calc_eom_static at iwb == 0 and
 * calc_eom_dmft).
 *   iq(npts): 1-based positions in q_data.
 *   chi_dens_out/chi_magn_out/bubble_nl_out(ndim^2,ndim^2,npts): the non-local chi parts and the
 *   non-local bubble, as chi_qw_dens/magn(:,:,iqw) and bubble_nl(:,:,iqw) of the reference.
 *   status(npts): per-point ADGA_* code (may be NULL).  A point with a non-zero status (singular or
 *   non-finite BSE matrix, or no convergence) is left out of the self-energy; its chi outputs are
 *   whatever the failed solve gave.  The other points of the batch are not affected.
 * Replaces: the loop body src/main.F90:635-781 (accumulate_chi0, the inline BSE zgemm calls,
 * inverse_matrix, calc_chi_qw, calc_eom_dynamic/_dmft/_static) for a batch of points.
 */
int adga_run_points(adga_ctx *ctx, int npts, const int *iq, int precision, double *chi_dens_out,
                    double *chi_magn_out, double *bubble_nl_out, int *status);

/*
 * Downloads the accumulated self-energies, with Sigma(-nu-1) = Sigma(nu)^H filled in:
 *   sigma_nl(ndim,ndim,2*iwfmax_small,nkp_eom), sigma_dmft(ndim,ndim,2*iwfmax_small),
 *   sigma_hf(ndim,ndim,nkp_eom) (zero without V(q)).  Any pointer may be NULL.
 * Replaces: reading sigma_nl/sigma_dmft/sigma_hf after the loop (src/main.F90:841-855).
 */
int adga_get_sigma(adga_ctx *ctx, double *sigma_nl, double *sigma_dmft, double *sigma_hf);

/*
 * Module-level entry points: one call per call of a reference module procedure, on the context's
 * one-particle data, for replacement Fortran modules with the reference's names and interfaces
 * (one_particle_quant_module, susc_module, eom_module) that leave the calling program unchanged.
 * Caller arrays are host memory in the reference's layouts; every call runs on the GPU and returns
 * when its outputs are written.  The batched path (adga_run_points) is the fast one: these cost a
 * host-device round trip per call.
 */

/*
 * chi0q(ndim^2, ndim^2, 2*iwfmax_small): the bubble of the q whose k-grid index is q_k (1-based) at
 * omega iwb, every nu of the box: -beta/nkp sum_k G(k, nu)_il G(k-q, nu-omega)_kj at
 * [(i-1)*ndim + j, (l-1)*ndim + k], i.e. the sum of accumulate_chi0 over ik divided by nkp.  With the
 * option bubble_range >= 1 every nu of the large range instead: 2*(iwmax - iwbmax_small) frequencies
 * from -iwmax + iwbmax_small (iwstart .. iwstop of the debug option "Bubble").
 * Replaces: accumulate_chi0 over the k loop (src/one_particle_quant_module.f90:228-290,
 * src/main.F90:647-653).
 */
int adga_bubble_q(adga_ctx *ctx, int iwb, int q_k, double *chi0q);

/*
 * Diagonals (ndim^2, 2*iwfmax_small) of chi0_loc(nu) = -beta giw(nu, i) giw(nu - omega, j) at
 * i1 = (i-1)*ndim + j and of its inverse, for every nu of the box (either pointer may be NULL); with
 * bubble_range >= 1 for every nu of the large range, as adga_bubble_q.
 * Replaces: get_chi0_loc, get_chi0_loc_inv (src/one_particle_quant_module.f90:192-225).
 */
int adga_chi0_loc(adga_ctx *ctx, int iwb, double *chi0w_diag, double *chi0w_inv_diag);

/*
 * chi_qw(ndim^2, ndim^2) += beta^-2 sum_{i3, nu} interm3(:, {i3, nu}) chi0_sum(i3, :, nu), with
 * interm3(ndim^2, maxdim) and chi0_sum(ndim^2, ndim^2, 2*iwfmax_small).
 * Replaces: calc_chi_qw (src/susc_module.f90:34-56).
 */
int adga_chi_qw(adga_ctx *ctx, const double *interm3, const double *chi0_sum, double *chi_qw);

/*
 * The dynamic EOM of one point (q list position iq, 1-based, at omega iwb): m_tot from eta_d, eta_m,
 * gamma^q_d (and V(q), gamma^w_d when v is not NULL), added into the caller's
 * sigma_nl(ndim, ndim, 2*iwfmax_small, nkp_eom) with Sigma(-nu-1) = Sigma(nu)^H.  k - q is computed
 * from the context's q and EOM k lists (the kq_ind_eom table index_kq_eom builds).
 * Replaces: calc_eom_dynamic (src/eom_module.f90:165-233).
 */
int adga_eom_dynamic(adga_ctx *ctx, int iwb, int iq, const double *etaqd, const double *etaqm,
                     const double *gammawd, const double *gammaqd, const double *v, double *sigma_nl);

/*
 * The local EOM with the DMFT g (default branch): sigma_dmft(ndim, ndim, 2*iwfmax_small) inout.
 * Replaces: calc_eom_dmft (src/eom_module.f90:115-161) without the debug option Dmft_phbar.
 */
int adga_eom_dmft(adga_ctx *ctx, int iwb, const double *gammawd, double *sigma_dmft);

/*
 * The static EOM terms of q list position iq: at iq == 1 the local Hartree term into
 * sigma_dmft(ndim, ndim, 2*iwfmax_small); with nonlocal and v(ndim^2, ndim^2) not NULL (do_vq) the
 * non-local Hartree term (iq == 1) and the Fock term into sigma_hf(ndim, ndim, nkp_eom).
 * Replaces: calc_eom_static (src/eom_module.f90:39-111).
 */
int adga_eom_static(adga_ctx *ctx, int iq, const double *v, int nonlocal, double *sigma_dmft, double *sigma_hf);

/*
 * gkiw(ndim, ndim) = G(k_ik, nu_iw) (ik 1-based, nu = pi (2 iw + 1) / beta) with Sigma = siw(iw, :)
 * on the diagonal when skiw is NULL (iw in [-iwmax, iwmax)), or the caller's skiw(ndim, ndim).
 * Replaces: get_gkiw(ik, iwf, iwb) with iw = iwf - iwb, get_gkiw_dga
 * (src/one_particle_quant_module.f90:81-127).
 */
int adga_get_gkiw(adga_ctx *ctx, int ik, int iw, const double *skiw, double *gkiw);

/*
 * n_dmft(ndim) = 2/beta sum_{iw >= 0} Re giw(iw, :) + 1/2 from the context's giw; it also becomes the
 * context's n_dmft.  Replaces: get_ndmft (src/one_particle_quant_module.f90:295-313).
 */
int adga_get_ndmft(adga_ctx *ctx, double *n_dmft);

/*
 * sigma_sum(ndim, ndim, 2*iwfmax_small, nkp_eom) += delta_ab siw(nu, a).
 * Replaces: add_siw_dmft (src/eom_module.f90:237-251).
 */
int adga_add_siw_dmft(adga_ctx *ctx, double *sigma_sum);

/*
 * Derivatives (tangent-linear form) of the per-point outputs and of Sigma_nl along ndir directions of the
 * inputs, for gradients with respect to model parameters; complex128 throughout, no V(q).  Per direction:
 *   dhk(ndim,ndim,nkp,ndir), dsiw/dgiw(2*iwmax,ndim,ndir), dmu(ndir) (real), du/du_tilde(ndim^2,ndim^2,ndir)
 * (dc is held fixed).  The derivative of the G table is formed at once (dG = G (dH + dSigma - dmu) G).
 * No counterpart in the reference (it has no gradients).
 */
int adga_set_tangents(adga_ctx *ctx, int ndir, const double *dhk, const double *dsiw, const double *dgiw,
                      const double *dmu, const double *du, const double *du_tilde);

/*
 * After adga_set_local_omega(_box) at an omega: the derivatives of its local data, given the derivatives
 * of chi_loc_dens/magn (N(omega), N(omega), ndir; NULL: zero).  Outputs (may be NULL), (ndim^2, ndim^2, ndir):
 * the derivatives of chi_loc_dens/magn_out and bubble_loc_out of adga_set_local_omega.
 */
int adga_set_local_omega_tangent(adga_ctx *ctx, const double *dchi_loc_dens, const double *dchi_loc_magn,
                                 double *dchi_loc_dens_out, double *dchi_loc_magn_out, double *dbubble_loc_out);

/*
 * adga_run_points in complex128 that also returns, per direction, the derivatives of chi_dens, chi_magn and
 * bubble_nl (ndim^2, ndim^2, npts, ndir; may be NULL) and accumulates the derivative of Sigma_nl on the
 * device.  Each direction reuses the batch's LU factors: a right-hand side and a solve; the bubble's and
 * the EOM's derivatives follow the context's "bubble" and "eom" options (k-grid convolutions: the bubble's
 * derivative for every q of the omega, dm accumulated on the q grid and convolved when the omega's EOM is
 * flushed; direct sums: per batch).
 */
int adga_run_points_tangent(adga_ctx *ctx, int npts, const int *iq, double *chi_dens_out, double *chi_magn_out,
                            double *bubble_nl_out, double *dchi_dens_out, double *dchi_magn_out,
                            double *dbubble_nl_out, int *status);

/* dsigma_nl(ndim, ndim, 2*iwfmax_small, nkp_eom, ndir): the accumulated derivative of sigma_nl (zeroed by
   adga_set_tangents and adga_reset_sigma). */
int adga_get_sigma_tangent(adga_ctx *ctx, double *dsigma_nl);

/* Zeroes the accumulated self-energies and drops the per-omega bubble cache (the next batch forms the
   bubble of its omega again). */
int adga_reset_sigma(adga_ctx *ctx);

/*
 * One-particle routines around the loop, on the context's one-particle data (hk, siw, dc, mu,
 * beta, k_data_eom).  G(k, nu)^-1 = (i nu + mu) - H(k) - dc(1,:) - Sigma(k, nu) on the diagonal and
 * -H(k) - Sigma(k, nu) off it, with nu = pi (2 iw + 1) / beta, iw = 0 .. iwmax-1.  All outputs are
 * complex(kind=8) arrays in the reference's layouts (first index fastest, frequency offset iwmax or
 * iwfmax_small); occupations are real-valued complex numbers as in the reference.
 */

/*
 * giw(2*iwmax, ndim) = (1/nkp) sum_k G_ii(k, nu) with Sigma = siw, giw(-nu-1) = giw(nu)^*.  As get_giw
 * overwrites the module's giw, the context's giw (the local bubble's input) becomes the result:
 * adga_set_local_omega must be called again before the next adga_run_points (ADGA_E_STATE otherwise).
 * Replaces: get_giw (src/one_particle_quant_module.f90:40-79).
 */
int adga_get_giw(adga_ctx *ctx, double *giw);

/*
 * n_fock(nkp, ndim, ndim) = 2/beta sum_{iw=0}^{iwmax-1} Re G(k, nu) + 1/2 on the diagonal, Sigma = siw.
 * Replaces: get_nfock (src/one_particle_quant_module.f90:337-361).
 */
int adga_get_nfock(adga_ctx *ctx, double *n_fock);

/*
 * sigma_sum(ndim, ndim, 2*iwfmax_small, nkp): the self-energy on the whole k grid after
 * add_siw_dmft (requires nkp_eom == nkp, as the reference's call does).  NULL: the context's own,
 * formed on the device from the Sigma_nl it has accumulated (as adga_get_sigma returns it) plus siw
 * on the orbital diagonal, without crossing PCIe (a single-process run's sigma_sum).  Outputs (may be NULL):
 * sigma_loc(ndim, ndim, 2*iwfmax_small) = its k mean; gloc(2*iwmax, ndim, ndim) = (1/nkp) sum_k G
 * with Sigma = sigma_sum inside the box (iw < iwfmax_small) and siw outside, gloc(-nu-1) = gloc(nu)^H.
 * Replaces: get_sigma_g_loc (src/one_particle_quant_module.f90:124-180).
 */
int adga_get_sigma_g_loc(adga_ctx *ctx, const double *sigma_sum, double *sigma_loc, double *gloc);

/*
 * sigma_sum(ndim, ndim, 2*iwfmax_small, nkp_eom) as above, on the EOM k list (NULL: the context's
 * own, as above).  Outputs (may be NULL):
 * n_dga_k(nkp_eom, ndim, ndim) = 2/beta sum_iw Re G(k_data_eom(ik), nu) + 1/2 on the diagonal
 * (Sigma = sigma_sum inside the box, siw outside); n_dga(ndim) = mean over ik of its diagonal.
 * Replaces: get_ndga (src/one_particle_quant_module.f90:363-393).
 */
int adga_get_ndga(adga_ctx *ctx, const double *sigma_sum, double *n_dga_k, double *n_dga);

/*
 * Statistics of the last adga_run_points call (nout <= 16 values): out[0] = max refinement
 * iterations, out[1] = max final relative residual of the points that succeeded, out[2..8] =
 * milliseconds spent in the reference's stage buckets Local, Chi0, Gammas, Inversion, Eta rest, EOM,
 * Chi (src/main.F90:484-783), out[9] = milliseconds of the batched LU factorisations, out[10] =
 * their number, out[11..13] = milliseconds, algorithmic flops and launches of the factorisation's
 * dominant GEMM kernel (the largest tile shape), out[14] = matrices solved again in complex128
 * after their complex64 refinement failed, out[15] = points with a non-zero status.
 */
int adga_get_stats(adga_ctx *ctx, double *out, int nout);

/* Copies the message of the last failure (NUL-terminated, truncated to len). */
int adga_last_error(adga_ctx *ctx, char *buf, int len);

int adga_ctx_destroy(adga_ctx *ctx);

/*
 * In-place inverse of a complex(kind=8) n x n column-major matrix on the GPU (LU with partial
 * pivoting, then n solves); ierr > 0 is the first zero pivot like zgetrf's info.  n above
 * ADGA_INVERSE_MAX_N returns ADGA_E_ARG.
 * Replaces: inverse_matrix_z (src/lapack_module.f90:40-70).
 */
int adga_inverse_matrix_z(int n, double *m, int *ierr);

/* largest n adga_inverse_matrix_z_noalloc takes */
#define ADGA_INVERSE_NOALLOC_MAX_N 1024

/*
 * adga_inverse_matrix_z without any device allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the matrix, its right-hand
 * sides, the panel inverses and the pivots live in one file-scope device array of the library
 * (34.6 MB, held by each GPU the library runs on until the process ends).  Same arguments, results and
 * errors as adga_inverse_matrix_z; n above ADGA_INVERSE_NOALLOC_MAX_N returns ADGA_E_ARG.  Every call on
 * a GPU shares that GPU's array: calls that overlap on one GPU race.
 * No counterpart in the reference.
 */
int adga_inverse_matrix_z_noalloc(int n, double *m, int *ierr);

/*
 * lambda correction of the final susceptibility (no counterpart in the reference, which has no
 * lambda correction).  The physical susceptibility of a point is the sum over i, j of the reduced
 * susceptibility's entries chi[(i,i), (j,j)] (those src/hdf5_module.f90:1479-1495 writes); the corrected
 * one is chi^lambda = 1/(1/chi + lambda); lambda is fixed by the sum rule
 *   Re S(lambda) = target,  S(lambda) = 1/(beta nqp) sum over every (q, omega) of chi^lambda,
 * with the local sum 1/beta sum_omega chi_loc as the usual target.  The sums run in FP64 on the device.
 */
typedef struct adga_lambda adga_lambda;

/*
 * chi_qw(ndim^2, ndim^2, nqp*(2*iwbmax+1)): the final susceptibility of every point of the omega-major
 * (omega, q) list (chi + bubble_nl + chi_loc + bubble_loc, the array main.F90 gathers as chi_qw_full,
 * src/main.F90:1047-1063); chi_loc(ndim^2, ndim^2, 2*iwbmax+1): the final local susceptibility
 * (chi_loc + bubble_loc, src/main.F90:935-936).  ADGA_E_ARG if no omega = 0 point has a positive
 * physical susceptibility (lambda then has no lower bound).
 */
int adga_lambda_create(int ndim, int nqp, int iwbmax, double beta, const double *chi_qw, const double *chi_loc,
                       adga_lambda **out);
/* S(lambda) and dS/dlambda = -1/(beta nqp) sum (chi^lambda)^2 (complex, interleaved) for nl lambdas */
int adga_lambda_sums(adga_lambda *h, int nl, const double *lambdas, double *s, double *ds);
/* target = Re 1/beta sum_omega chi_loc_phys(omega), the local sum rule */
int adga_lambda_target(adga_lambda *h, double *target);
/*
 * lambda in (-1/max Re chi_phys(q, omega = 0), inf) with |Re S(lambda) - target| <= tol max(1, |target|)
 * (Newton's method on the device sums, bisection where a step leaves the bracket); ADGA_E_NOCONV after
 * maxit steps, ADGA_E_NONFINITE when a sum is not finite, ADGA_E_ARG for a non-finite target.  iters
 * (may be NULL): the steps taken.
 */
int adga_lambda_solve(adga_lambda *h, double target, double tol, int maxit, double *lambda, int *iters);
/* chi_phys^lambda of every point, (nqp*(2*iwbmax+1)) complex */
int adga_lambda_chi(adga_lambda *h, double lambda, double *chi_phys);
int adga_lambda_last_error(adga_lambda *h, char *buf, int len);
int adga_lambda_destroy(adga_lambda *h);

#ifdef __cplusplus
}
#endif

#endif /* ADGA_AMD_H */
