// Host-side context of the C ABI (include/adga_amd.h) and the helpers its entry points share:
// HIP error checks that throw inside `guarded`, which turns them into status codes at the boundary.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <exception>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "adga_amd.h"
#include "adga_kernels.h"

using adga::c32;
using adga::Geom;
using adga::z64;

constexpr int ADGA_NSTATS = 16;  // values adga_get_stats reports

// LU lane counts tried per batch shape, and the factorisation times (ms) each gave
constexpr int LANE_CANDIDATES = 3;
struct LaneTune {
  float ms[LANE_CANDIDATES] = {};
  int issued = 0;  // candidates handed out for a timed trial so far (lane counts 1 .. issued)
  int done = 0;    // bit j: the time of lane count j + 1 is in ms[j]
  int best = 1;    // the fastest, once every candidate is timed
};

// Buffers of the FP64 re-solve of points whose complex64 refinement did not converge: the points'
// S and right-hand sides are gathered into these, solved in complex128 and the solutions scattered back.
struct FallbackBufs {
  int cap = 0;  // points
  z64 *S = nullptr, *B = nullptr, *Y = nullptr, *Z = nullptr, *R = nullptr, *M = nullptr, *inv = nullptr;
  z64 *W = nullptr, *Q = nullptr;  // V(q) rows and products of the gathered points
  double* rel = nullptr;
  int *piv = nullptr, *info = nullptr, *sel = nullptr;
};

struct adga_ctx {
  Geom g{};   // the context's geometry (the largest BSE box; the self-energies' box)
  Geom gw{};  // the current omega's geometry: g with the BSE box of that omega
  Geom eom_g{};  // geometry of the omega whose EOM is pending
  hipStream_t stream = nullptr;
  adga::LuLanes lanes;                     // streams 1.. of the LU lanes (default priority) and their events
  int lu_split = 0;                        // option "lu_streams": parts each LU batch is split into (0: tuned)
  // lu_streams = 0: per device and batch shape (precision, BSE size, matrices) the first factorisations
  // in the process run with 1 .. LANE_CANDIDATES lanes, each timed by ev_lu, and the later ones with the
  // fastest; every lane count runs the same kernels on the same data, so results do not depend on it
  int device = 0;      // the device current at creation: the context's streams belong to it
  long lane_key = -1;  // shape of this context's factorisation whose trial time is pending (-1: none)
  int lane_trial = 0;  // its lane count
  hipEvent_t ev[7] = {}, ev_lu[2] = {};    // stage and LU timing of adga_run_points
  std::string err;
  int refine_maxit = 12;
  double refine_tol = 1e-12;
  bool fallback = true;  // option "fallback": unconverged complex64 solves are solved again in complex128
  int batch = 256;
  int ldm = 0;
  // one-particle data (device)
  z64 *hk = nullptr, *siw = nullptr, *giw = nullptr, *dc = nullptr, *u = nullptr, *ut = nullptr, *ndmft = nullptr;
  z64* gt = nullptr;
  int* keom = nullptr;
  std::vector<int> q0;  // 0-based k index of each q (host)
  // non-local interaction V(q) (adga_set_vq): (n2, n2, nqp) on the device, n_fock for its Fock term,
  // the batch's q list positions, V rows W and n2 x n2 products
  z64 *vq = nullptr, *nfock = nullptr, *W = nullptr, *vqQ = nullptr;
  int* qpos = nullptr;
  int vq_cap = 0;
  // current bosonic frequency
  bool have_omega = false;
  int iwb = 0;
  z64 *xd = nullptr, *xm = nullptr, *cd = nullptr, *cm = nullptr, *gamd = nullptr, *gamm = nullptr;
  z64 *opgd = nullptr, *opgm = nullptr, *winv = nullptr, *chi0w = nullptr;
  z64 *chiloc_d = nullptr, *chiloc_m = nullptr, *bubloc = nullptr;
  // batch workspace
  int cap = 0;
  z64 *chi0q = nullptr, *S = nullptr, *gamq = nullptr, *Bp = nullptr, *Y = nullptr, *Z = nullptr, *R = nullptr,
      *mt = nullptr, *chi_out = nullptr, *bub_out = nullptr;
  void* M = nullptr;
  void* inv = nullptr;  // per-matrix panel inverses of the LU
  double* rel = nullptr;
  int *piv = nullptr, *info = nullptr, *qidx = nullptr;
  bool q_distinct = false;          // the current batch's q indices are pairwise different
  std::vector<unsigned char> qmark;  // host scratch of that test, zero between batches (nkp entries)
  int *act = nullptr, *keep = nullptr;  // per-matrix refinement flags, per-point "enters the self-energy" flags
  FallbackBufs fb;
  // LU operation lists (factor + solve of the appended right-hand sides; solve only), per precision
  std::map<int, std::vector<int>> ops_f[2];  // LU operation lists (host) per precision and BSE size
  std::map<int, std::vector<int>> ops_r[2];  // the same with the n^2 right-hand sides carried along
  // accumulated self-energies
  z64 *sig_nl = nullptr, *sig_dmft = nullptr, *sig_hf = nullptr;
  // k-grid convolutions: twiddles, Ghat table, per-omega bubble cache, pending EOM of one omega
  int bubble_mode = 2;  // 0 direct, 1 DFT, 2 DFT when the batch has >= 64 points or the cache is valid
  int eom_mode = 1;     // 0 direct, 1 DFT (accumulated per omega, flushed on omega change / get_sigma)
  bool do_eom = true;   // option "do_eom": the reference's do_eom (off for a q path, main.F90:341-347)
  z64 *tw = nullptr, *ghat = nullptr, *bub_p = nullptr, *bub_all = nullptr, *macc = nullptr, *mwork = nullptr;
  bool bub_valid = false;
  int bub_iwb = 0, bub_iwf = 0;
  // option "bubble_range": 0 the box, 1 every frequency the 1P data allow (the reference's debug option
  // "Bubble"; the G table then covers [-iwmax, iwmax)), 2 that and the Matsubara tail in bubble_loc
  int bubble_range = 0;
  z64 *outer_p = nullptr, *outer_all = nullptr, *bubout = nullptr;  // outer-frequency bubble of all q, one omega
  bool outer_valid = false;
  int outer_iwb = 0, outer_iwf = 0;
  bool eom_pending = false;
  int eom_iwb = 0;
  double stats[ADGA_NSTATS] = {0};  // see adga_get_stats
  // tangent directions (adga_set_tangents): derivative of the G table, of giw, mu, U and U~ per direction
  int ndir = 0;
  z64 *dgt = nullptr, *dgiw = nullptr, *du = nullptr, *dut = nullptr;
  double* dmu = nullptr;
  // derivatives of the current omega's local data, per direction (adga_set_local_omega_tangent)
  bool t_omega = false;
  int t_iwb = 0, t_iwf = 0;
  z64 *t_dw = nullptr, *t_dwinv = nullptr, *t_dcd = nullptr, *t_dcm = nullptr, *t_dgamd = nullptr, *t_dgamm = nullptr;
  // per-batch derivative workspace and the derivative of Sigma_nl per direction
  int t_cap = 0;
  z64 *t_dchi0q = nullptr, *t_dS = nullptr, *t_Z1 = nullptr, *t_Z2 = nullptr, *t_B = nullptr, *t_dY = nullptr,
      *t_dgamq = nullptr, *t_dmt = nullptr, *t_dchi = nullptr, *t_dbub = nullptr;
  z64* t_sig = nullptr;
  // EOM derivatives by k-grid convolution (eom mode 1): dm per direction on the q grid, the transform of
  // each direction's dG table, scratch
  z64 *t_macc = nullptr, *t_dghat = nullptr, *t_mwork = nullptr;
  // bubble derivatives of every q of one omega (k-grid convolutions), per direction, and their scratch
  z64 *t_dbub_all = nullptr, *t_bub_p = nullptr;
  bool t_bub_valid = false;
  int t_bub_iwb = 0, t_bub_iwf = 0;
  // device staging of caller arrays for the module-level entry points (api_modules.cpp)
  z64* stage[6] = {};
  size_t stage_n[6] = {};
  adga::LuTimer lut;       // events around the dominant GEMM kernel's launches
};

// derivative outputs of adga_run_points_tangent: (ndim^2, ndim^2, npts, ndir) each, may be NULL
struct TangentOut {
  double *dchid, *dchim, *dbub;
};

namespace adga_host {

struct HipError {
  hipError_t e;
  const char* what;
};

#define CK(x)                                   \
  do {                                          \
    hipError_t e_ = (x);                        \
    if (e_ != hipSuccess) throw HipError{e_, #x}; \
  } while (0)

template <typename F>
inline int guarded(adga_ctx* c, F&& f) {
  try {
    return f();
  } catch (const HipError& he) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)he.e, hipGetErrorString(he.e), he.what);
    if (c) c->err = buf;
    return he.e == hipErrorOutOfMemory ? ADGA_E_NOMEM : ADGA_E_HIP;
  } catch (const std::bad_alloc&) {
    if (c) c->err = "host allocation failed";
    return ADGA_E_NOMEM;
  } catch (const std::exception& ex) {
    if (c) c->err = std::string("host exception: ") + ex.what();
    return ADGA_E_HIP;
  } catch (...) {
    if (c) c->err = "unexpected host exception";
    return ADGA_E_HIP;
  }
}

template <typename T>
void dalloc(T** p, size_t count) {
  CK(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}
inline void upload(z64* d, const double* h, size_t count, hipStream_t s) {
  CK(hipMemcpyAsync(d, h, count * sizeof(z64), hipMemcpyHostToDevice, s));
}

inline int fail(adga_ctx* c, int code, const char* msg) {
  if (c) c->err = msg;
  return code;
}

// device buffer released on every exit path
template <typename T = z64>
struct DevBuf {
  T* p = nullptr;
  explicit DevBuf(size_t count) { dalloc(&p, count); }
  ~DevBuf() { dfree(p); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};
// n_fock of the context's one-particle data (get_nfock), formed on first use
inline void ensure_nfock(adga_ctx* c) {
  if (c->nfock) return;
  dalloc(&c->nfock, (size_t)c->g.nkp * c->g.n2);
  if (!adga::launch_op_nfock(c->g, c->hk, c->siw, c->dc, c->nfock, c->stream))
    throw HipError{hipErrorInvalidValue, "n_fock: orbital count above 8"};
}

// the context's geometry with the "box" set to the bubble's frequency range: the vertex box, or with the
// option bubble_range >= 1 every frequency for which G(nu - omega) exists, [-iwmax + iwbmax, iwmax - iwbmax)
// (iwstart/iwstop of the debug option "Bubble", config_module.f90:467-477)
inline Geom bubble_geom(const adga_ctx* c) {
  Geom g = c->g;
  if (c->bubble_range >= 1) {
    g.iwf = g.iwmax - g.iwbmax;
    g.nf = 2 * g.iwf;
    g.N = g.n2 * g.nf;
  }
  return g;
}

inline int op_finish(adga_ctx* c, bool launched, const char* what) {
  if (!launched) {
    c->err = std::string(what) + ": orbital count above 8";
    return ADGA_E_ARG;
  }
  CK(hipGetLastError());
  return ADGA_OK;
}

// derivatives of one batch of points (positions p0 .. p0 + nb - 1 of npts) along every tangent direction,
// after the forward work of the batch (api_tangent.cpp)
void tangent_batch(adga_ctx* c, int nb, int p0, int npts, const std::vector<int>& keep, int nfailed,
                   const TangentOut& out);
void free_tangent(adga_ctx* c);
// the twiddles and the k-grid transform of the G table (formed once); the pending omega's EOM flushed
// (with its derivatives when tangent directions are set) (api.cpp)
void conv_tables(adga_ctx* c);
void flush_eom(adga_ctx* c);

}  // namespace adga_host
