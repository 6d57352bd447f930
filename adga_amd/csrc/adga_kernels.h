// Host-side launchers of the gfx950 kernels (internal to libadga_amd.so).
#pragma once
#include <vector>

#include "adga_common.h"

namespace adga {

void launch_gtable(const Geom& g, const z64* hk, const z64* siw, const z64* dc, z64* gt, hipStream_t s);
// one-particle routines (onep.hip); false for an orbital count the kernels do not cover (n > 8)
bool launch_op_giw(const Geom& g, const z64* hk, const z64* siw, const z64* dc, z64* giw, hipStream_t s);
bool launch_op_nfock(const Geom& g, const z64* hk, const z64* siw, const z64* dc, z64* nfock, hipStream_t s);
size_t op_sigma_loc_scratch(const Geom& g);  // complex elements of the scratch launch_op_sigma_g_loc needs
bool launch_op_sigma_g_loc(const Geom& g, const z64* hk, const z64* siw, const z64* dc, const z64* ssum, z64* sloc,
                           z64* gloc, z64* scratch, hipStream_t s);
bool launch_op_ndga(const Geom& g, const z64* hk, const z64* siw, const z64* dc, const int* keom, const z64* ssum,
                    z64* ndk, z64* nd, hipStream_t s);
void launch_local(const Geom& g, int iwb, const z64* giw, const z64* xd, const z64* xm, z64* winv, z64* chi0w,
                  z64* cd, z64* cm, z64* gam_d, z64* gam_m, z64* opg_d, z64* opg_m, z64* chiloc_d,
                  z64* chiloc_m, z64* bubloc, hipStream_t s);
// gt2 (may be null: gt) is the table of G(K - q, nu - omega); accumulate adds into chi0q
void launch_bubble(const Geom& g, int iwb, int npts, const int* qidx, const z64* gt, z64* chi0q, hipStream_t s,
                   const z64* gt2 = nullptr, bool accumulate = false);
// Out = C_ch^T Z_ch for the channels ch0 .. ch0 + nch - 1 (C row-major as the left operand, Z and Out in the
// residual layouts), plus B - Y when B is given, plus Out itself when accumulate
void resid_gemm(const Geom& g, int npts, int ch0, int nch, const z64* cd, const z64* cm, const z64* Z, z64* Out,
                const z64* B, const z64* Y, hipStream_t s, bool accumulate = false);
void launch_smat(const Geom& g, int npts, const z64* chi0q, const z64* winv, z64* S, hipStream_t s);
template <typename T>
void launch_build(const Geom& g, int npts, int ldm, const z64* S, const z64* cd, const z64* cm, cplx<T>* M,
                  hipStream_t s);
void launch_gamq(const Geom& g, int npts, const z64* S, const z64* cd, z64* Z, z64* G2, hipStream_t s);
// the right operand of gamma^q_d's product (S of each point in the residual product's layout)
void launch_gamq_z(const Geom& g, int npts, const z64* S, z64* Z, hipStream_t s);
size_t lu_inv_elems(int N);
std::vector<int> lu_ops_factor(int N, int W);
// factor + forward elimination of the right-hand-side columns [N, W) only (no backward substitution)
std::vector<int> lu_ops_factor_fwd(int N, int W);
// the same list as lu_ops_factor written into buf (at most cap ints); returns the number of ints the list needs
size_t lu_ops_factor_into(int N, int W, int* buf, size_t cap);
// Events around the launches of the largest GEMM tile shape (the factorisation's dominant kernel):
// mark() records one event; collect() (after the stream has completed) returns the summed
// milliseconds of the bracketed launches and resets.
struct LuTimer {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  double flops = 0.0;  // algorithmic flops of the bracketed launches (8 m n K each)
  int launches = 0;
  void mark(hipStream_t s);
  double collect();
  ~LuTimer();
};
// streams an LU batch may be split over: lane 0 is the launch stream itself, lanes 1 .. n-1 the streams
// s[1..], with the events that order them (all streams at the default priority)
constexpr int LU_MAX_LANES = 4;
struct LuLanes {
  int n = 1;
  hipStream_t s[LU_MAX_LANES] = {};
  hipEvent_t lag[LU_MAX_LANES] = {}, join[LU_MAX_LANES] = {};
};
template <typename T>
bool launch_lu(int nmat, int N, int W, int ldm, const int* ops, size_t nops, cplx<T>* M, int* piv, cplx<T>* inv,
               int* info, hipStream_t s, LuTimer* timer = nullptr, const LuLanes* lanes = nullptr);
template <typename T>
inline bool launch_lu(int nmat, int N, int W, int ldm, const std::vector<int>& ops, cplx<T>* M, int* piv,
                      cplx<T>* inv, int* info, hipStream_t s, LuTimer* timer = nullptr,
                      const LuLanes* lanes = nullptr) {
  return launch_lu<T>(nmat, N, W, ldm, ops.data(), ops.size(), M, piv, inv, info, s, timer, lanes);
}
// `active` (may be null): per-matrix flags; matrices with a zero flag are skipped (Y left as is).
// from_m: only the backward substitution, of the right-hand sides the factorisation carried along in the
// columns [N, N + nrhs) of M (lu_ops_factor_fwd); R is then not read
template <typename T>
bool launch_lu_solve(int nmat, int N, int nrhs, int ldm, const cplx<T>* M, const int* perm, const cplx<T>* inv,
                     const z64* R, z64* Y, bool accumulate, hipStream_t s, const int* active = nullptr,
                     bool from_m = false);
// whether the nrhs columns of a solution fit one workgroup's LDS at once (elem bytes per complex element):
// else every solve streams the factors once per column group
bool lu_solve_one_group(int N, int nrhs, size_t elem);
// the right-hand sides B ([matrix][row][nrhs], FP64) into the columns [N, N + nrhs) of each matrix of M
template <typename T>
void launch_rhs_into_m(int nmat, int N, int nrhs, int ldm, const z64* B, cplx<T>* M, hipStream_t s);
// The non-local interaction's rank-n2 term of the density BSE matrix (main.F90:710-719): W holds, per
// point, the N x n2 rows (2/beta^2) chi0^q(nu) V(q); opgd = 1 + gamma^w_d; Q is n2 x n2 per point of scratch.
struct VqTerm {
  const z64* W;
  const z64* opgd;
  z64* Q;
};
void launch_rhs(const Geom& g, int npts, const z64* S, const z64* cd, const z64* cm, const z64* opgd,
                const z64* opgm, z64* Z, z64* B, hipStream_t s, const VqTerm* vq = nullptr);
void launch_residual(const Geom& g, int npts, const z64* S, const z64* cd, const z64* cm, const z64* B,
                     const z64* Y, z64* Z, z64* R, double* rel, hipStream_t s, const VqTerm* vq = nullptr);
// V(q) (vq.hip): W of a batch (qpos: 0-based q list positions, vq: (n2, n2, nqp)); the density matrices'
// term; Out_{2b} += (1 + gamma^w_d)^T (W^T X) with X = (1 + gamma^w_d)^T (Y null) or the density solution Y
void launch_vq_w(const Geom& g, int npts, const int* qpos, const z64* vq, const z64* chi0q, z64* W, hipStream_t s);
template <typename T>
void launch_vq_build(const Geom& g, int npts, int ldm, const VqTerm& v, cplx<T>* M, hipStream_t s);
void launch_vq_add(const Geom& g, int npts, const VqTerm& v, const z64* Y, z64* Out, hipStream_t s);
void launch_eom_fock(const Geom& g, int npts, const int* qpos, const int* qidx, const int* keom, const z64* vq,
                     const z64* nfock, z64* shf, hipStream_t s);
void launch_eom_hartree_v(const Geom& g, const z64* v, const z64* ndmft, z64* shf, hipStream_t s);
// chi_ch(q) and bubble_nl of a batch; eta = the batch's solutions Y (2 npts matrices, eta's layout)
void launch_chi(const Geom& g, int npts, const z64* eta, const z64* chi0q, const z64* chi0w, const z64* gamd,
                const z64* gamm, const z64* bubloc, z64* chi_out, z64* bub_out, hipStream_t s);
// zeroes the per-point blocks (`len` complex values each) of the points whose flag is zero
void launch_zero_points(int npts, long len, const int* keep, z64* a, hipStream_t s);
// dst[i*len .. ) = src[sel[i]*len .. ) for i < k (gather) or the reverse (scatter)
void launch_gather(int k, long len, const int* sel, const z64* src, z64* dst, bool scatter, hipStream_t s);
void launch_eom_local(const Geom& g, int iwb, bool do_static, const z64* u, const z64* gamd, const z64* giw,
                      const z64* ndmft, z64* sdmft, hipStream_t s);
void launch_fill_neg(const Geom& g, int nk, z64* sigma, hipStream_t s);
// the local Hartree term alone (eom_module.f90:52-70)
void launch_eom_static_local(const Geom& g, const z64* u, const z64* ndmft, z64* sdmft, hipStream_t s);
// module-level kernels (modules.hip)
void launch_chi0_loc_diag(const Geom& g, int iwb, const z64* giw, z64* w, z64* winv, hipStream_t s);
bool launch_gkiw(const Geom& g, int k, int iw, const z64* hk, const z64* siw, const z64* dc, const z64* skiw,
                 z64* out, hipStream_t s);
void launch_chi_qw(const Geom& g, const z64* X, const z64* Y, z64* chi, hipStream_t s);
void launch_ndmft(const Geom& g, const z64* giw, z64* nd, hipStream_t s);
void launch_add_siw(const Geom& g, int nk, const z64* siw, z64* ss, hipStream_t s);
// with vq non-null: V(q_b) = vq + qpos[b] n4 enters the density eta term and -V gamma^w_d / (beta nqp) is added
void launch_mtot(const Geom& g, int npts, const z64* u, const z64* ut, const z64* eta, const z64* gamq, z64* mt,
                 hipStream_t s, const int* qpos = nullptr, const z64* vq = nullptr, const z64* gamd = nullptr,
                 bool accumulate = false);
void launch_eom_direct(const Geom& g, int iwb, int npts, const int* qidx, const int* keom, const z64* mt,
                       const z64* gt, z64* sigma, hipStream_t s);
// tangent-linear kernels (tangent.hip)
bool launch_dgtable(const Geom& g, int ndir, const z64* gt, const z64* dhk, const z64* dsiw, const double* dmu,
                    z64* dgt, hipStream_t s);
void launch_local_tangent(const Geom& g, int iwb, const z64* giw, const z64* dgiw, const z64* xd, const z64* xm,
                          const z64* dxd, const z64* dxm, const z64* winv, const z64* gamd, const z64* gamm,
                          const z64* chi0w, z64* dw, z64* dwinv, z64* dcd, z64* dcm, z64* dgamd, z64* dgamm,
                          z64* dchiloc_d, z64* dchiloc_m, z64* dbubloc, hipStream_t s);
void launch_dsmat(const Geom& g, int npts, const z64* chi0q, const z64* dchi0q, const z64* winv, const z64* dwinv,
                  z64* dS, hipStream_t s);
void launch_tangent_rhs(const Geom& g, int npts, const z64* S, const z64* dS, const z64* cd, const z64* cm,
                        const z64* dcd, const z64* dcm, const z64* opgd, const z64* opgm, const z64* dopgd,
                        const z64* dopgm, const z64* Y, z64* Z1, z64* Z2, z64* B, hipStream_t s);
void launch_dchi(const Geom& g, int npts, const z64* chi0q, const z64* dchi0q, const z64* chi0w, const z64* dchi0w,
                 const z64* gamd, const z64* gamm, const z64* dgamd, const z64* dgamm, const z64* Y, const z64* dY,
                 z64* dchi_out, z64* dbub_out, hipStream_t s);
// k-grid convolutions (kconv.hip)
void dft3(const Geom& g, long nfield, int sign, const z64* tw, z64* a, z64* b, hipStream_t s);
void launch_ghat(const Geom& g, const z64* tw, const z64* gt, z64* tmp, z64* ghat, hipStream_t s);
void launch_bubble_all(const Geom& g, int iwb, const z64* tw, const z64* ghat, z64* P, z64* out, hipStream_t s);
void launch_bubble_gather(const Geom& g, int npts, const int* qidx, const z64* all, z64* chi0q, hipStream_t s);
// the bubble's derivative for every q of an omega from the transforms of G and dG (product rule)
void launch_bubble_all_tangent(const Geom& g, int iwb, const z64* tw, const z64* ghat, const z64* dghat, z64* P,
                               z64* out, hipStream_t s);
// distinct: the batch's q indices are pairwise different (then one addition per element runs in parallel
// over the points; else the points are added in order by one thread per element)
void launch_eom_scatter(const Geom& g, int npts, const int* qidx, const z64* mt, z64* macc, hipStream_t s,
                        bool distinct = false);
// derivatives carried by the EOM flush: per direction d, the accumulator of dm_d on the q grid (macc + d*nm),
// the k-grid transform of dG_d (dghat + d*ng) and d Sigma_nl (sigma + d*nsig); work: nm scratch
struct EomTangent {
  int ndir = 0;
  z64* macc = nullptr;
  const z64* dghat = nullptr;
  z64* work = nullptr;
  z64* sigma = nullptr;
  size_t nm = 0, ng = 0, nsig = 0;
};
void launch_eom_flush(const Geom& g, int iwb, const int* keom, const z64* tw, const z64* ghat, z64* macc, z64* work,
                      z64* sigma, hipStream_t s, const EomTangent* tan = nullptr);
// bubble beyond the box: large range and Matsubara tail (tail.hip)
void launch_bubloc_range(const Geom& g, int iwb, int lo, int hi, const z64* giw, z64* bubloc, hipStream_t s);
void launch_tail(const Geom& g, int iwb, int K, const z64* giw, const z64* bubloc, z64* out, hipStream_t s);
void launch_bubble_outer(const Geom& g, int iwb, int lo, int hi, const z64* tw, const z64* ghat, z64* P, z64* out,
                         hipStream_t s);
void launch_bubble_outer_add(const Geom& g, int npts, const int* qidx, const z64* all, z64* bub_out, hipStream_t s);
// lambda correction sums (lambda.hip)
void launch_chi_phys(int n, long npts, const z64* chi, z64* out, hipStream_t s);
void launch_lambda_sums(long npts, int nl, int nblk, double scale, const double* lams, const z64* chi, z64* part,
                        z64* out, hipStream_t s);
void launch_lambda_apply(long npts, double lam, const z64* chi, z64* out, hipStream_t s);

}  // namespace adga
