// C ABI (include/adga_amd.h) over the gfx950 kernels: context, device buffers, batching,
// refinement loop and error reporting.  No exception crosses the extern "C" boundary.
#include <cmath>
#include <cstring>
#include <mutex>

#include "ctx.h"

using namespace adga_host;

namespace {

void free_batch(adga_ctx* c) {
  dfree(c->chi0q); dfree(c->S); dfree(c->gamq); dfree(c->Bp); dfree(c->Y); dfree(c->Z); dfree(c->R);
  dfree(c->mt); dfree(c->chi_out); dfree(c->bub_out); dfree(c->M); dfree(c->inv); dfree(c->rel);
  dfree(c->piv); dfree(c->info); dfree(c->qidx); dfree(c->act); dfree(c->keep);
  c->cap = 0;
}

void free_fallback(adga_ctx* c) {
  FallbackBufs& f = c->fb;
  dfree(f.S); dfree(f.B); dfree(f.Y); dfree(f.Z); dfree(f.R); dfree(f.M); dfree(f.inv); dfree(f.rel);
  dfree(f.W); dfree(f.Q);
  dfree(f.piv); dfree(f.info); dfree(f.sel);
  f.cap = 0;
}

void ensure_batch(adga_ctx* c, int nb) {
  if (nb <= c->cap) return;
  free_batch(c);
  const Geom& g = c->g;
  size_t per_pt = (size_t)g.nf * g.n4;
  size_t pmat = (size_t)g.N * g.n2;
  dalloc(&c->chi0q, nb * per_pt);
  dalloc(&c->S, nb * per_pt);
  dalloc(&c->gamq, 2 * nb * pmat);
  dalloc(&c->Bp, 2 * nb * pmat);
  dalloc(&c->Y, 2 * nb * pmat);
  dalloc(&c->Z, 2 * nb * pmat);
  dalloc(&c->R, 2 * nb * pmat);
  dalloc(&c->mt, nb * pmat);
  dalloc(&c->chi_out, 2 * nb * (size_t)g.n4);
  dalloc(&c->bub_out, nb * (size_t)g.n4);
  dalloc((char**)&c->M, 2 * nb * (size_t)g.N * c->ldm * sizeof(z64));
  dalloc((char**)&c->inv, 2 * nb * adga::lu_inv_elems(g.N) * sizeof(z64));
  dalloc(&c->rel, 2 * nb);
  dalloc(&c->piv, 2 * nb * (size_t)g.N);
  dalloc(&c->info, 2 * nb);
  dalloc(&c->qidx, nb);
  dalloc(&c->act, 2 * nb);
  dalloc(&c->keep, nb);
  c->cap = nb;
}

// V(q) workspace of a batch of nb points
void ensure_vq_batch(adga_ctx* c, int nb) {
  if (nb <= c->vq_cap) return;
  dfree(c->W); dfree(c->vqQ); dfree(c->qpos);
  const Geom& g = c->g;
  dalloc(&c->W, (size_t)nb * g.N * g.n2);
  dalloc(&c->vqQ, (size_t)nb * g.n4);
  dalloc(&c->qpos, nb);
  c->vq_cap = nb;
}

void ensure_conv(adga_ctx* c) {
  const Geom& g = c->g;
  if (c->tw) return;
  // forward twiddles e^{-2 pi i R k / n} of the three axes
  std::vector<z64> tw;
  for (int n : {g.nkx, g.nky, g.nkz})
    for (int R = 0; R < n; ++R)
      for (int k = 0; k < n; ++k) {
        double ph = -2.0 * M_PI * (double)((long)R * k % n) / n;
        tw.push_back(adga::mk(cos(ph), sin(ph)));
      }
  dalloc(&c->tw, tw.size());
  CK(hipMemcpy(c->tw, tw.data(), sizeof(z64) * tw.size(), hipMemcpyHostToDevice));
  size_t ng = (size_t)g.nw * g.n2 * g.nkp;
  z64* tmp = nullptr;
  dalloc(&tmp, ng);
  dalloc(&c->ghat, ng);
  adga::launch_ghat(g, c->tw, c->gt, tmp, c->ghat, c->stream);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(c->stream));
  dfree(tmp);
}

void eom_flush(adga_ctx* c);

// the G table and its k-grid transform over every frequency of the 1P data, nu in [-iwmax, iwmax), for the
// large bubble range (option "bubble_range")
void widen_gtable(adga_ctx* c) {
  eom_flush(c);  // a pending EOM reads the transform of the current table
  const int iwmax = c->g.iwmax;
  for (Geom* g : {&c->g, &c->gw, &c->eom_g}) {
    g->glo = -iwmax;
    g->nw = 2 * iwmax;
  }
  dfree(c->gt);
  dalloc(&c->gt, (size_t)c->g.nw * c->g.nkp * c->g.n2);
  adga::launch_gtable(c->g, c->hk, c->siw, c->dc, c->gt, c->stream);
  CK(hipGetLastError());
  dfree(c->ghat);
  dfree(c->tw);  // both formed again by ensure_conv
  c->bub_valid = false;
  c->outer_valid = false;
  CK(hipStreamSynchronize(c->stream));
}

void eom_flush(adga_ctx* c) {
  if (!c->eom_pending) return;
  const size_t nm = (size_t)c->g.iwf * c->g.n4 * c->g.nkp;
  adga::EomTangent tan;
  if (c->ndir > 0 && c->t_macc) {
    tan.ndir = c->ndir;
    tan.macc = c->t_macc;
    tan.dghat = c->t_dghat;
    tan.work = c->t_mwork;
    tan.sigma = c->t_sig;
    tan.nm = nm;
    tan.ng = (size_t)c->g.nw * c->g.n2 * c->g.nkp;
    tan.nsig = (size_t)c->g.n2 * c->g.nf * c->g.nkp_eom;
  }
  adga::launch_eom_flush(c->eom_g, c->eom_iwb, c->keom, c->tw, c->ghat, c->macc, c->mwork, c->sig_nl, c->stream,
                         tan.ndir ? &tan : nullptr);
  CK(hipGetLastError());
  CK(hipMemsetAsync(c->macc, 0, sizeof(z64) * nm, c->stream));
  if (tan.ndir) CK(hipMemsetAsync(c->t_macc, 0, sizeof(z64) * nm * tan.ndir, c->stream));
  c->eom_pending = false;
}

// LU operation list of the current omega's BSE size; prec 0: complex128 LU, prec 1: complex64 LU
const std::vector<int>& ensure_ops(adga_ctx* c, int prec) {
  const int N = c->gw.N;
  std::vector<int>& ops = c->ops_f[prec][N];
  if (ops.empty()) ops = adga::lu_ops_factor(N, N);
  return ops;
}

// whether the current omega's factorisation carries the n^2 right-hand sides as columns [N, N + n^2) of each
// matrix: their forward elimination then rides along in the factorisation's updates, and the first solve is
// only the backward substitution.  Chosen where the solution's columns do not fit the LDS at once, so that
// every solve pass streams the factors once per column group (at N = 4000: +7% points/s refined); where
// they fit, the extra columns cost the factorisation about what the forward pass saves (measured at
// N = 1080: 143 -> 148 ms per refined factorisation of 4096 matrices), and they are left out.
bool carry_rhs(const adga_ctx* c, int prec) {
  const Geom& g = c->gw;
  return !adga::lu_solve_one_group(g.N, g.n2, prec == 0 ? sizeof(z64) : sizeof(c32)) && c->ldm >= g.N + g.n2;
}

// The lane tuning table of the process, per (device, batch shape).  The fastest lane count is a property
// of the device and the shape, so every context shares it; contexts may run on several host threads.
std::mutex g_lane_mu;
std::map<std::pair<int, long>, LaneTune> g_lane_tune;

void lane_result(LaneTune& lt, int lanes, float ms) {
  const int bit = 1 << (lanes - 1);
  if (lt.done & bit) return;
  lt.ms[lanes - 1] = ms;
  lt.done |= bit;
  if (lt.done == (1 << LANE_CANDIDATES) - 1)
    lt.best = 1 + (int)(std::min_element(lt.ms, lt.ms + LANE_CANDIDATES) - lt.ms);
}

// lane count of the next factorisation of nmat matrices of size N.  A fixed option value is used as is.
// With the tuned default the best lane count depends on the batch shape (measured: 2 lanes fastest for
// 4096 matrices of N = 1080, 1 lane for 256 of N = 2000 .. 4000, 3 for 8000 of N = 1080, by 0.5-2%), so
// a shape's first factorisations on the device each take one untried candidate; the context keeps the
// trial in lane_key / lane_trial and lane_store files its time when the batch is done.  A trial left
// pending by a failed batch counts as the slowest.
int pick_lanes(adga_ctx* c, int prec, int N, int nmat) {
  std::lock_guard<std::mutex> lk(g_lane_mu);
  if (c->lane_key >= 0) lane_result(g_lane_tune[{c->device, c->lane_key}], c->lane_trial, INFINITY);
  c->lane_key = -1;
  if (c->lu_split > 0) return c->lu_split;
  const long key = ((long)prec << 44) | ((long)N << 22) | (long)nmat;  // nmat <= 2^21, N <= 4096
  LaneTune& lt = g_lane_tune[{c->device, key}];
  if (lt.done == (1 << LANE_CANDIDATES) - 1) return lt.best;
  if (lt.issued < LANE_CANDIDATES) {
    c->lane_key = key;
    c->lane_trial = ++lt.issued;
    return c->lane_trial;
  }
  // every candidate is out, some still running in another context: the fastest timed so far
  int best = 1;
  float bm = INFINITY;
  for (int j = 0; j < LANE_CANDIDATES; ++j)
    if ((lt.done >> j & 1) && lt.ms[j] < bm) {
      bm = lt.ms[j];
      best = j + 1;
    }
  return best;
}

// files the time of the context's pending trial
void lane_store(adga_ctx* c, float ms) {
  if (c->lane_key < 0) return;
  std::lock_guard<std::mutex> lk(g_lane_mu);
  lane_result(g_lane_tune[{c->device, c->lane_key}], c->lane_trial, ms);
  c->lane_key = -1;
}

const std::vector<int>& ensure_ops_rhs(adga_ctx* c, int prec) {
  const int N = c->gw.N;
  std::vector<int>& ops = c->ops_r[prec][N];
  if (ops.empty()) ops = adga::lu_ops_factor_fwd(N, N + c->gw.n2);
  return ops;
}

// FP64 re-solve of the points `pts` (batch positions) whose complex64 refinement did not converge:
// their S and right-hand sides are gathered, the density and magnetic matrices rebuilt in complex128,
// factored, solved and checked, and the solutions scattered back into the batch's Y.  Updates the
// points' entries of relh / infoh and returns the number of matrices re-solved.
int fallback_fp64(adga_ctx* c, const std::vector<int>& pts, std::vector<double>& relh, std::vector<int>& infoh) {
  const Geom& g = c->gw;
  hipStream_t s = c->stream;
  if (pts.empty()) return 0;
  const std::vector<int>& ops64 = ensure_ops(c, 0);
  FallbackBufs& f = c->fb;
  if (f.cap == 0) {
    // sized for the largest box (c->g); at most 4 GB of complex128 matrices (two per point), at most
    // 32 points per round
    const Geom& G = c->g;
    const size_t per_mat = (size_t)G.N * c->ldm;
    const size_t fit = ((size_t)4 << 30) / (2 * per_mat * sizeof(z64));
    const int cap = (int)std::max<size_t>(1, std::min<size_t>(32, fit));
    const size_t pmat = (size_t)G.N * G.n2;
    dalloc(&f.S, (size_t)cap * G.nf * G.n4);
    dalloc(&f.B, 2 * cap * pmat);
    dalloc(&f.Y, 2 * cap * pmat);
    dalloc(&f.Z, 2 * cap * pmat);
    dalloc(&f.R, 2 * cap * pmat);
    dalloc(&f.M, 2 * cap * per_mat);
    dalloc(&f.inv, 2 * cap * adga::lu_inv_elems(G.N));
    dalloc(&f.rel, 2 * cap);
    dalloc(&f.piv, 2 * cap * (size_t)G.N);
    dalloc(&f.info, 2 * cap);
    dalloc(&f.sel, cap);
    f.cap = cap;
  }
  if (c->vq && !f.W) {
    dalloc(&f.W, (size_t)f.cap * c->g.N * c->g.n2);
    dalloc(&f.Q, (size_t)f.cap * c->g.n4);
  }
  const long slen = (long)g.nf * g.n4, blen = 2L * g.N * g.n2;
  std::vector<double> r;
  std::vector<int> inf;
  int done = 0;
  for (size_t i0 = 0; i0 < pts.size(); i0 += f.cap) {
    const int k = (int)std::min<size_t>(f.cap, pts.size() - i0);
    CK(hipMemcpyAsync(f.sel, &pts[i0], sizeof(int) * k, hipMemcpyHostToDevice, s));
    adga::launch_gather(k, slen, f.sel, c->S, f.S, false, s);
    adga::launch_gather(k, blen, f.sel, c->Bp, f.B, false, s);
    adga::launch_build<double>(g, k, c->ldm, f.S, c->cd, c->cm, f.M, s);
    const adga::VqTerm fv{f.W, c->opgd, f.Q};
    if (c->vq) {
      adga::launch_gather(k, (long)g.N * g.n2, f.sel, c->W, f.W, false, s);
      adga::launch_vq_build<double>(g, k, c->ldm, fv, f.M, s);
    }
    if (!adga::launch_lu<double>(2 * k, g.N, g.N, c->ldm, ops64, f.M, f.piv, f.inv, f.info, s) ||
        !adga::launch_lu_solve<double>(2 * k, g.N, g.n2, c->ldm, f.M, f.piv, f.inv, f.B, f.Y, false, s))
      throw HipError{hipErrorInvalidValue, "FP64 re-solve: BSE size outside the LU kernel's range"};
    adga::launch_residual(g, k, f.S, c->cd, c->cm, f.B, f.Y, f.Z, f.R, f.rel, s, c->vq ? &fv : nullptr);
    adga::launch_gather(k, blen, f.sel, f.Y, c->Y, true, s);
    r.resize(2 * k);
    inf.resize(2 * k);
    CK(hipMemcpyAsync(r.data(), f.rel, sizeof(double) * 2 * k, hipMemcpyDeviceToHost, s));
    CK(hipMemcpyAsync(inf.data(), f.info, sizeof(int) * 2 * k, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (int j = 0; j < k; ++j)
      for (int ch = 0; ch < 2; ++ch) {
        relh[2 * pts[i0 + j] + ch] = r[2 * j + ch];
        infoh[2 * pts[i0 + j] + ch] = inf[2 * j + ch];
      }
    done += 2 * k;
  }
  return done;
}

// FP64 iterative refinement of the complex64 solves of a batch of nb points (2 nb matrices), per
// matrix: a matrix stops when its relative residual max|r|/max|b| reaches refine_tol; it is handed to
// the FP64 re-solve when its complex64 LU met an exactly zero pivot, a residual is not finite, a
// correction did not at least halve its residual (stagnation or divergence: the complex64 factors
// cannot precondition it, roughly cond(A) 2^-24 >= 1/2) or the iteration limit is reached.  The
// corrections run only on the matrices still refining.  On return relh / infoh hold each matrix's
// final residual and pivot status and resolved[m] = 1 for the matrices solved again in complex128
// (both of a point; none when the option "fallback" is off); returns the number of corrections of
// the slowest matrix.
// infoh must have been queued for download (the complex64 LU's zero-pivot flags) on the stream.
int refine(adga_ctx* c, int nb, std::vector<double>& relh, std::vector<int>& infoh, std::vector<int>& resolved) {
  const Geom& g = c->gw;
  hipStream_t s = c->stream;
  const int nmat = 2 * nb;
  enum { ACTIVE = 0, DONE = 1, FP64 = 2 };
  std::vector<int> state(nmat, ACTIVE), act(nmat);
  std::vector<double> prev(nmat, INFINITY);
  int iters = 0;
  const adga::VqTerm vt{c->W, c->opgd, c->vqQ};
  for (int it = 0;; ++it) {
    adga::launch_residual(g, nb, c->S, c->cd, c->cm, c->Bp, c->Y, c->Z, c->R, c->rel, s, c->vq ? &vt : nullptr);
    CK(hipMemcpyAsync(relh.data(), c->rel, sizeof(double) * nmat, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    int nact = 0;
    for (int m = 0; m < nmat; ++m) {
      if (state[m] != ACTIVE) continue;
      const double r = relh[m];
      if (infoh[m] != 0 || !std::isfinite(r)) state[m] = FP64;
      else if (r <= c->refine_tol) state[m] = DONE;
      else if (it >= c->refine_maxit || (it > 0 && r > 0.5 * prev[m])) state[m] = FP64;
      else ++nact;
      prev[m] = r;
    }
    if (nact == 0) break;
    for (int m = 0; m < nmat; ++m) act[m] = state[m] == ACTIVE;
    CK(hipMemcpyAsync(c->act, act.data(), sizeof(int) * nmat, hipMemcpyHostToDevice, s));
    // correction: solve with the complex64 factors for the (rounded) residual and add it to the FP64
    // solution
    adga::launch_lu_solve<float>(nmat, g.N, g.n2, c->ldm, (const c32*)c->M, c->piv, (const c32*)c->inv, c->R, c->Y,
                                 true, s, c->act);
    iters = it + 1;
  }
  std::vector<int> pts;
  resolved.assign(nmat, 0);
  for (int b = 0; b < nb && c->fallback; ++b)
    if (state[2 * b] == FP64 || state[2 * b + 1] == FP64) {
      pts.push_back(b);
      resolved[2 * b] = resolved[2 * b + 1] = 1;
    }
  c->stats[14] += fallback_fp64(c, pts, relh, infoh);
  return iters;
}

}  // namespace

namespace adga_host {
void conv_tables(adga_ctx* c) { ensure_conv(c); }
void flush_eom(adga_ctx* c) { eom_flush(c); }
}  // namespace adga_host

extern "C" {

int adga_ctx_create(const adga_dims* d, const int* q_data, const int* k_data_eom, const double* hk,
                    const double* siw, const double* giw, const double* dc, const double* u, const double* u_tilde,
                    const double* n_dmft, adga_ctx** out) {
  if (!out || !d || !q_data || !k_data_eom || !hk || !siw || !giw || !dc || !u || !u_tilde || !n_dmft)
    return ADGA_E_ARG;
  *out = nullptr;
  if (d->ndim < 1 || d->ndim > 5 || d->nkpx < 1 || d->nkpy < 1 || d->nkpz < 1 || d->iwfmax_small < 1 ||
      d->iwbmax_small < 0 || d->nqp < 1 || d->nkp_eom < 1 || !(d->beta > 0))
    return ADGA_E_ARG;
  if (d->iwmax < d->iwfmax_small + d->iwbmax_small) return ADGA_E_ARG;  // config_module.f90:600-607
  adga_ctx* c = new (std::nothrow) adga_ctx();
  if (!c) return ADGA_E_NOMEM;
  int rc = guarded(c, [&]() -> int {
    Geom& g = c->g;
    g.n = d->ndim;
    g.n2 = g.n * g.n;
    g.n4 = g.n2 * g.n2;
    g.iwf = d->iwfmax_small;
    g.iws = g.iwf;
    g.nf = 2 * g.iwf;
    g.iwbmax = d->iwbmax_small;
    g.N = g.n2 * g.nf;
    g.nkx = d->nkpx; g.nky = d->nkpy; g.nkz = d->nkpz;
    g.nkp = g.nkx * g.nky * g.nkz;
    g.glo = -g.iwf - g.iwbmax;
    g.nw = g.nf + 2 * g.iwbmax;
    g.iwmax = d->iwmax;
    g.nqp = d->nqp;
    g.nkp_eom = d->nkp_eom;
    g.beta = d->beta;
    g.mu = d->mu;
    // N columns, and the n^2 right-hand sides when the factorisation carries them (carry_rhs)
    c->ldm = (g.N + (adga::lu_solve_one_group(g.N, g.n2, sizeof(z64)) ? 0 : g.n2) + 15) / 16 * 16;
    c->gw = g;
    c->eom_g = g;
    c->q0.resize(g.nqp);
    for (int i = 0; i < g.nqp; ++i) {
      if (q_data[i] < 1 || q_data[i] > g.nkp) throw HipError{hipErrorInvalidValue, "q_data out of range"};
      c->q0[i] = q_data[i] - 1;
    }
    for (int i = 0; i < g.nkp_eom; ++i)
      if (k_data_eom[i] < 1 || k_data_eom[i] > g.nkp) throw HipError{hipErrorInvalidValue, "k_data_eom out of range"};
    CK(hipGetDevice(&c->device));
    CK(hipStreamCreate(&c->stream));
    c->lanes.s[0] = c->stream;
    for (int j = 1; j < adga::LU_MAX_LANES; ++j) CK(hipStreamCreate(&c->lanes.s[j]));
    for (int j = 0; j < adga::LU_MAX_LANES; ++j) {
      CK(hipEventCreateWithFlags(&c->lanes.lag[j], hipEventDisableTiming));
      CK(hipEventCreateWithFlags(&c->lanes.join[j], hipEventDisableTiming));
    }
    for (auto& e : c->ev) CK(hipEventCreate(&e));
    for (auto& e : c->ev_lu) CK(hipEventCreate(&e));
    hipStream_t s = c->stream;
    dalloc(&c->hk, (size_t)g.n2 * g.nkp);
    dalloc(&c->siw, (size_t)2 * g.iwmax * g.n);
    dalloc(&c->giw, (size_t)2 * g.iwmax * g.n);
    dalloc(&c->dc, (size_t)2 * g.n);
    dalloc(&c->u, (size_t)g.n4);
    dalloc(&c->ut, (size_t)g.n4);
    dalloc(&c->ndmft, (size_t)g.n);
    dalloc(&c->gt, (size_t)g.nw * g.nkp * g.n2);
    dalloc(&c->keom, (size_t)g.nkp_eom);
    upload(c->hk, hk, (size_t)g.n2 * g.nkp, s);
    upload(c->siw, siw, (size_t)2 * g.iwmax * g.n, s);
    upload(c->giw, giw, (size_t)2 * g.iwmax * g.n, s);
    upload(c->dc, dc, (size_t)2 * g.n, s);
    upload(c->u, u, (size_t)g.n4, s);
    upload(c->ut, u_tilde, (size_t)g.n4, s);
    upload(c->ndmft, n_dmft, (size_t)g.n, s);
    CK(hipMemcpyAsync(c->keom, k_data_eom, sizeof(int) * g.nkp_eom, hipMemcpyHostToDevice, s));
    adga::launch_gtable(g, c->hk, c->siw, c->dc, c->gt, s);
    CK(hipGetLastError());
    size_t NN = (size_t)g.N * g.N, pm = (size_t)g.n2 * g.N;
    dalloc(&c->xd, NN); dalloc(&c->xm, NN); dalloc(&c->cd, NN); dalloc(&c->cm, NN);
    dalloc(&c->gamd, pm); dalloc(&c->gamm, pm); dalloc(&c->opgd, pm); dalloc(&c->opgm, pm);
    dalloc(&c->winv, (size_t)g.n2 * g.nf); dalloc(&c->chi0w, (size_t)g.n2 * g.nf);
    dalloc(&c->chiloc_d, (size_t)g.n4); dalloc(&c->chiloc_m, (size_t)g.n4); dalloc(&c->bubloc, (size_t)g.n4);
    dalloc(&c->sig_nl, (size_t)g.n2 * g.nf * g.nkp_eom);
    dalloc(&c->sig_dmft, (size_t)g.n2 * g.nf);
    dalloc(&c->sig_hf, (size_t)g.n2 * g.nkp_eom);
    CK(hipMemsetAsync(c->sig_hf, 0, sizeof(z64) * g.n2 * g.nkp_eom, s));
    CK(hipMemsetAsync(c->sig_nl, 0, sizeof(z64) * g.n2 * g.nf * g.nkp_eom, s));
    CK(hipMemsetAsync(c->sig_dmft, 0, sizeof(z64) * g.n2 * g.nf, s));
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
  if (rc != ADGA_OK) {
    std::string msg = c->err;
    adga_ctx_destroy(c);
    (void)msg;
    return rc;
  }
  *out = c;
  return ADGA_OK;
}

int adga_set_option(adga_ctx* c, const char* key, double v) {
  if (!c || !key) return ADGA_E_ARG;
  if (!strcmp(key, "batch")) {
    if (!(v >= 1 && v <= 1 << 20)) return fail(c, ADGA_E_ARG, "batch must be in [1, 2^20]");
    c->batch = (int)v;
    return ADGA_OK;
  }
  if (!strcmp(key, "refine_tol")) {
    if (!(v > 0 && v < 1)) return fail(c, ADGA_E_ARG, "refine_tol must be in (0, 1)");
    c->refine_tol = v;
    return ADGA_OK;
  }
  if (!strcmp(key, "lu_streams")) {
    if (!(v >= 0 && v <= adga::LU_MAX_LANES && v == (int)v))
      return fail(c, ADGA_E_ARG, "lu_streams must be 0 (tuned per batch shape) or 1 .. 4");
    c->lu_split = (int)v;
    return ADGA_OK;
  }
  if (!strcmp(key, "refine_maxit")) {
    if (!(v >= 1 && v <= 100)) return fail(c, ADGA_E_ARG, "refine_maxit must be in [1, 100]");
    c->refine_maxit = (int)v;
    return ADGA_OK;
  }
  if (!strcmp(key, "bubble")) {
    if (!(v == 0 || v == 1 || v == 2)) return fail(c, ADGA_E_ARG, "bubble must be 0 (direct), 1 (dft) or 2 (auto)");
    c->bubble_mode = (int)v;
    return ADGA_OK;
  }
  if (!strcmp(key, "bubble_range")) {
    if (!(v == 0 || v == 1 || v == 2))
      return fail(c, ADGA_E_ARG, "bubble_range must be 0 (box), 1 (large range) or 2 (large range and tail)");
    if (v >= 1 && c->ndir > 0) return fail(c, ADGA_E_STATE, "bubble_range 1 or 2 has no tangent-linear form");
    return guarded(c, [&]() -> int {
      if (v >= 1 && c->g.glo != -c->g.iwmax) widen_gtable(c);
      if ((int)v != c->bubble_range) c->have_omega = false;  // bubble_loc of the current omega is stale
      c->bubble_range = (int)v;
      return ADGA_OK;
    });
  }
  if (!strcmp(key, "fallback")) {
    if (!(v == 0 || v == 1)) return fail(c, ADGA_E_ARG, "fallback must be 0 or 1");
    c->fallback = v == 1;
    return ADGA_OK;
  }
  if (!strcmp(key, "do_eom")) {
    if (!(v == 0 || v == 1)) return fail(c, ADGA_E_ARG, "do_eom must be 0 or 1");
    if ((v == 1) != c->do_eom && c->eom_pending)
      return fail(c, ADGA_E_STATE, "do_eom cannot change while an omega is pending (call adga_get_sigma)");
    c->do_eom = v == 1;
    return ADGA_OK;
  }
  if (!strcmp(key, "eom")) {
    if (!(v == 0 || v == 1)) return fail(c, ADGA_E_ARG, "eom must be 0 (direct) or 1 (dft)");
    if ((int)v != c->eom_mode && c->eom_pending)
      return fail(c, ADGA_E_STATE, "eom mode cannot change while an omega is pending (call adga_get_sigma)");
    c->eom_mode = (int)v;
    return ADGA_OK;
  }
  return fail(c, ADGA_E_ARG, "unknown option");
}

int adga_set_vq(adga_ctx* c, const double* v_q) {
  if (!c) return ADGA_E_ARG;
  if (c->eom_pending) return fail(c, ADGA_E_STATE, "V(q) cannot change while an omega is pending (call adga_get_sigma)");
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    if (!v_q) {
      dfree(c->vq);
      return ADGA_OK;
    }
    const size_t nv = (size_t)g.n4 * g.nqp;
    if (!c->vq) dalloc(&c->vq, nv);
    upload(c->vq, v_q, nv, c->stream);
    CK(hipStreamSynchronize(c->stream));
    return ADGA_OK;
  });
}

int adga_set_local_omega_box(adga_ctx* c, int iwb, int iwf_box, const double* xd, const double* xm, double* chid_out,
                             double* chim_out, double* bubloc_out) {
  if (!c || !xd || !xm) return ADGA_E_ARG;
  if (iwb < -c->g.iwbmax || iwb > c->g.iwbmax) return fail(c, ADGA_E_ARG, "iwb outside [-iwbmax_small, iwbmax_small]");
  if (iwf_box < 1 || iwf_box > c->g.iwf) return fail(c, ADGA_E_ARG, "box outside [1, iwfmax_small]");
  return guarded(c, [&]() -> int {
    hipStream_t s = c->stream;
    if (c->eom_pending && (c->eom_iwb != iwb || c->eom_g.iwf != iwf_box)) eom_flush(c);
    if (c->bub_iwb != iwb || c->bub_iwf != iwf_box) c->bub_valid = false;
    if (c->outer_iwb != iwb || c->outer_iwf != iwf_box) c->outer_valid = false;
    Geom g = c->g;  // the box of this omega: N = n^2 * 2 * iwf_box
    g.iwf = iwf_box;
    g.nf = 2 * iwf_box;
    g.N = g.n2 * g.nf;
    size_t NN = (size_t)g.N * g.N;
    upload(c->xd, xd, NN, s);
    upload(c->xm, xm, NN, s);
    adga::launch_local(g, iwb, c->giw, c->xd, c->xm, c->winv, c->chi0w, c->cd, c->cm, c->gamd, c->gamm, c->opgd,
                       c->opgm, c->chiloc_d, c->chiloc_m, c->bubloc, s);
    // bubble_loc over the large range [-K, K-1] (main.F90:513-520 with the debug option "Bubble"); it is
    // also what bubble_nl subtracts, while the tail enters the returned bubble_loc only
    const int K = c->g.iwmax - c->g.iwbmax;
    if (c->bubble_range >= 1) adga::launch_bubloc_range(g, iwb, -K, K - 1, c->giw, c->bubloc, s);
    const z64* bl = c->bubloc;
    if (c->bubble_range == 2) {
      if (!c->bubout) dalloc(&c->bubout, (size_t)g.n4);
      adga::launch_tail(g, iwb, K, c->giw, c->bubloc, c->bubout, s);
      bl = c->bubout;
    }
    CK(hipGetLastError());
    if (chid_out) CK(hipMemcpyAsync(chid_out, c->chiloc_d, sizeof(z64) * g.n4, hipMemcpyDeviceToHost, s));
    if (chim_out) CK(hipMemcpyAsync(chim_out, c->chiloc_m, sizeof(z64) * g.n4, hipMemcpyDeviceToHost, s));
    if (bubloc_out) CK(hipMemcpyAsync(bubloc_out, bl, sizeof(z64) * g.n4, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    c->gw = g;
    c->iwb = iwb;
    c->have_omega = true;
    return ADGA_OK;
  });
}

int adga_set_local_omega(adga_ctx* c, int iwb, const double* xd, const double* xm, double* chid_out,
                         double* chim_out, double* bubloc_out) {
  if (!c) return ADGA_E_ARG;
  return adga_set_local_omega_box(c, iwb, c->g.iwf, xd, xm, chid_out, chim_out, bubloc_out);
}

}  // extern "C"

namespace {
// the per-point work of adga_run_points; with `tan` (adga_run_points_tangent) also the derivatives of every
// batch along the context's tangent directions
int run_points_impl(adga_ctx* c, int npts, const int* iq, int precision, double* chid, double* chim, double* bubnl,
                    int* status, const TangentOut* tan) {
  if (precision != ADGA_PREC_FP64 && precision != ADGA_PREC_REFINED) return fail(c, ADGA_E_ARG, "bad precision");
  if (!c->have_omega) return fail(c, ADGA_E_STATE, "adga_set_local_omega must be called before adga_run_points");
  for (int p = 0; p < npts; ++p)
    if (iq[p] < 1 || iq[p] > c->g.nqp) return fail(c, ADGA_E_ARG, "iq outside [1, nqp]");
  return guarded(c, [&]() -> int {
    const Geom& g = c->gw;  // the BSE box of the current omega
    hipStream_t s = c->stream;
    const int n2 = g.n2;
    std::fill(c->stats, c->stats + ADGA_NSTATS, 0.0);
    c->lut.used = 0;
    c->lut.flops = 0.0;
    int worst = ADGA_OK;
    hipEvent_t* ev = c->ev;
    std::vector<int> qh, qp, infoh, st, keep, resolved;
    std::vector<double> relh;
    std::vector<z64> chih, bubh;
    for (int p0 = 0; p0 < npts; p0 += c->batch) {
      const int nb = std::min(c->batch, npts - p0);
      ensure_batch(c, nb);
      qh.resize(nb);
      for (int b = 0; b < nb; ++b) qh[b] = c->q0[iq[p0 + b] - 1];
      CK(hipMemcpyAsync(c->qidx, qh.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
      // whether every q of the batch is different (the EOM scatter then runs over the points in parallel)
      c->qmark.resize(c->g.nkp);
      c->q_distinct = true;
      for (int b = 0; b < nb; ++b) {
        if (c->qmark[qh[b]]) c->q_distinct = false;
        c->qmark[qh[b]] = 1;
      }
      for (int b = 0; b < nb; ++b) c->qmark[qh[b]] = 0;
      if (c->vq) {
        ensure_vq_batch(c, nb);
        qp.resize(nb);
        for (int b = 0; b < nb; ++b) qp[b] = iq[p0 + b] - 1;
        CK(hipMemcpyAsync(c->qpos, qp.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
      }
      const adga::VqTerm vt{c->W, c->opgd, c->vqQ};
      const adga::VqTerm* vtp = c->vq ? &vt : nullptr;
      const int nmat = 2 * nb;
      CK(hipEventRecord(ev[0], s));
      // local EOM for points at the Gamma point (main.F90:635-639)
      for (int b = 0; b < nb && c->do_eom; ++b)
        if (iq[p0 + b] == 1) {
          adga::launch_eom_local(g, c->iwb, c->iwb == 0, c->u, c->gamd, c->giw, c->ndmft, c->sig_dmft, s);
          if (c->vq && c->iwb == 0) adga::launch_eom_hartree_v(g, c->vq, c->ndmft, c->sig_hf, s);
        }
      if (c->vq && c->do_eom && c->iwb == 0) {
        // non-local Fock term of every point at omega = 0 (calc_eom_static, main.F90:635-637)
        ensure_nfock(c);
        adga::launch_eom_fock(g, nb, c->qpos, c->qidx, c->keom, c->vq, c->nfock, c->sig_hf, s);
      }
      CK(hipEventRecord(ev[1], s));
      const bool bub_hit = c->bub_valid && c->bub_iwb == c->iwb && c->bub_iwf == g.iwf;
      const bool use_dft = c->bubble_mode == 1 || (c->bubble_mode == 2 && (bub_hit || nb >= 64));
      if (use_dft) {
        ensure_conv(c);
        if (!bub_hit) {
          size_t nb_all = (size_t)c->g.nf * g.n4 * g.nkp;  // sized for the largest box
          if (!c->bub_p) {
            dalloc(&c->bub_p, nb_all);
            dalloc(&c->bub_all, nb_all);
          }
          adga::launch_bubble_all(g, c->iwb, c->tw, c->ghat, c->bub_p, c->bub_all, s);
          c->bub_valid = true;
          c->bub_iwb = c->iwb;
          c->bub_iwf = g.iwf;
        }
        adga::launch_bubble_gather(g, nb, c->qidx, c->bub_all, c->chi0q, s);
      } else {
        adga::launch_bubble(g, c->iwb, nb, c->qidx, c->gt, c->chi0q, s);
      }
      adga::launch_smat(g, nb, c->chi0q, c->winv, c->S, s);
      if (vtp) adga::launch_vq_w(g, nb, c->qpos, c->vq, c->chi0q, c->W, s);
      adga::launch_gamq(g, nb, c->S, c->cd, c->Z, c->gamq, s);
      adga::launch_rhs(g, nb, c->S, c->cd, c->cm, c->opgd, c->opgm, c->Z, c->Bp, s, vtp);
      CK(hipEventRecord(ev[2], s));
      int iters = 0;
      double maxrel = 0.0;
      bool lu_ok;
      const int prec = precision == ADGA_PREC_FP64 ? 0 : 1;
      const bool rhs = carry_rhs(c, prec);
      const std::vector<int>& ops = rhs ? ensure_ops_rhs(c, prec) : ensure_ops(c, prec);
      const int W = rhs ? g.N + n2 : g.N;
      adga::LuLanes lanes = c->lanes;
      lanes.n = pick_lanes(c, prec, g.N, nmat);
      relh.assign(nmat, 0.0);
      infoh.assign(nmat, 0);
      resolved.assign(nmat, 0);
      if (precision == ADGA_PREC_FP64) {
        adga::launch_build<double>(g, nb, c->ldm, c->S, c->cd, c->cm, (z64*)c->M, s);
        if (vtp) adga::launch_vq_build<double>(g, nb, c->ldm, vt, (z64*)c->M, s);
        if (rhs) adga::launch_rhs_into_m<double>(nmat, g.N, n2, c->ldm, c->Bp, (z64*)c->M, s);
        CK(hipEventRecord(ev[3], s));
        CK(hipEventRecord(c->ev_lu[0], s));
        lu_ok = adga::launch_lu<double>(nmat, g.N, W, c->ldm, ops, (z64*)c->M, c->piv, (z64*)c->inv, c->info, s,
                                        &c->lut, &lanes);
        CK(hipEventRecord(c->ev_lu[1], s));
        if (lu_ok)
          lu_ok = adga::launch_lu_solve<double>(nmat, g.N, n2, c->ldm, (const z64*)c->M, c->piv, (const z64*)c->inv,
                                                rhs ? nullptr : c->Bp, c->Y, false, s, nullptr, rhs);
        if (lu_ok) {
          // one FP64 residual of the solution: reports the achieved accuracy and non-finite solves
          adga::launch_residual(g, nb, c->S, c->cd, c->cm, c->Bp, c->Y, c->Z, c->R, c->rel, s, vtp);
          CK(hipMemcpyAsync(relh.data(), c->rel, sizeof(double) * nmat, hipMemcpyDeviceToHost, s));
          CK(hipMemcpyAsync(infoh.data(), c->info, sizeof(int) * nmat, hipMemcpyDeviceToHost, s));
          CK(hipStreamSynchronize(s));
        }
      } else {
        adga::launch_build<float>(g, nb, c->ldm, c->S, c->cd, c->cm, (c32*)c->M, s);
        if (vtp) adga::launch_vq_build<float>(g, nb, c->ldm, vt, (c32*)c->M, s);
        if (rhs) adga::launch_rhs_into_m<float>(nmat, g.N, n2, c->ldm, c->Bp, (c32*)c->M, s);
        CK(hipEventRecord(ev[3], s));
        CK(hipEventRecord(c->ev_lu[0], s));
        lu_ok = adga::launch_lu<float>(nmat, g.N, W, c->ldm, ops, (c32*)c->M, c->piv, (c32*)c->inv, c->info, s,
                                       &c->lut, &lanes);
        CK(hipEventRecord(c->ev_lu[1], s));
        if (lu_ok)
          lu_ok = adga::launch_lu_solve<float>(nmat, g.N, n2, c->ldm, (const c32*)c->M, c->piv, (const c32*)c->inv,
                                               rhs ? nullptr : c->Bp, c->Y, false, s, nullptr, rhs);
        if (lu_ok) {
          CK(hipMemcpyAsync(infoh.data(), c->info, sizeof(int) * nmat, hipMemcpyDeviceToHost, s));
          iters = refine(c, nb, relh, infoh, resolved);
        }
      }
      if (!lu_ok) return fail(c, ADGA_E_ARG, "BSE size outside the supported range of the LU kernel");
      CK(hipGetLastError());
      // per-point status; points that failed are left out of the self-energy
      st.assign(nb, ADGA_OK);
      keep.assign(nb, 1);
      int nfailed = 0;
      for (int b = 0; b < nb; ++b) {
        for (int ch = 0; ch < 2; ++ch) {
          const int m = 2 * b + ch;
          int sm = ADGA_OK;
          if (infoh[m] != 0) sm = ADGA_E_SINGULAR;
          else if (!std::isfinite(relh[m])) sm = ADGA_E_NONFINITE;
          else if (precision == ADGA_PREC_REFINED && !resolved[m] && relh[m] > c->refine_tol) sm = ADGA_E_NOCONV;
          if (sm == ADGA_OK) maxrel = std::max(maxrel, relh[m]);
          if (st[b] == ADGA_OK) st[b] = sm;
        }
        if (st[b] != ADGA_OK) {
          keep[b] = 0;
          ++nfailed;
        }
      }
      CK(hipEventRecord(ev[4], s));
      adga::launch_chi(g, nb, c->Y, c->chi0q, c->chi0w, c->gamd, c->gamm, c->bubloc, c->chi_out, c->bub_out, s);
      if (c->bubble_range >= 1) {
        // the large range's frequencies outside the box enter bubble_nl (main.F90:644-663, debug option
        // "Bubble"): all q of the omega as one convolution, kept for the omega's next batches
        ensure_conv(c);
        if (!(c->outer_valid && c->outer_iwb == c->iwb && c->outer_iwf == g.iwf)) {
          if (!c->outer_p) {
            dalloc(&c->outer_p, (size_t)g.n4 * g.nkp);
            dalloc(&c->outer_all, (size_t)g.n4 * g.nkp);
          }
          const int K = c->g.iwmax - c->g.iwbmax;
          adga::launch_bubble_outer(g, c->iwb, -K, K - 1, c->tw, c->ghat, c->outer_p, c->outer_all, s);
          c->outer_valid = true;
          c->outer_iwb = c->iwb;
          c->outer_iwf = g.iwf;
        }
        adga::launch_bubble_outer_add(g, nb, c->qidx, c->outer_all, c->bub_out, s);
      }
      CK(hipEventRecord(ev[5], s));
      if (c->do_eom) {
        adga::launch_mtot(g, nb, c->u, c->ut, c->Y, c->gamq, c->mt, s, c->qpos, c->vq, c->gamd);  // eta = Y
        if (nfailed > 0) {
          CK(hipMemcpyAsync(c->keep, keep.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
          adga::launch_zero_points(nb, (long)n2 * g.N, c->keep, c->mt, s);
        }
        if (c->eom_mode == 1) {
          ensure_conv(c);
          if (!c->macc) {
            size_t nm = (size_t)c->g.iwf * g.n4 * g.nkp;  // sized for the largest box
            dalloc(&c->macc, nm);
            dalloc(&c->mwork, nm);
            CK(hipMemsetAsync(c->macc, 0, sizeof(z64) * nm, s));
          }
          if (c->eom_pending && (c->eom_iwb != c->iwb || c->eom_g.iwf != g.iwf)) eom_flush(c);
          adga::launch_eom_scatter(g, nb, c->qidx, c->mt, c->macc, s, c->q_distinct);
          c->eom_pending = true;
          c->eom_iwb = c->iwb;
          c->eom_g = g;
        } else {
          adga::launch_eom_direct(g, c->iwb, nb, c->qidx, c->keom, c->mt, c->gt, c->sig_nl, s);
        }
      }
      CK(hipEventRecord(ev[6], s));
      CK(hipGetLastError());
      if (tan) tangent_batch(c, nb, p0, npts, keep, nfailed, *tan);
      chih.resize((size_t)2 * nb * g.n4);
      bubh.resize((size_t)nb * g.n4);
      CK(hipMemcpyAsync(chih.data(), c->chi_out, sizeof(z64) * chih.size(), hipMemcpyDeviceToHost, s));
      CK(hipMemcpyAsync(bubh.data(), c->bub_out, sizeof(z64) * bubh.size(), hipMemcpyDeviceToHost, s));
      CK(hipStreamSynchronize(s));
      float ms[6];
      for (int k = 0; k < 6; ++k) CK(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
      c->stats[2 + 5] += ms[0];            // EOM (local part)
      c->stats[2 + 1] += ms[1];            // Chi0
      c->stats[2 + 2] += ms[2];            // Gammas (BSE build)
      c->stats[2 + 3] += ms[3];            // Inversion (LU, solve, refinement, FP64 re-solves)
      c->stats[2 + 6] += ms[4];            // Chi (eta + contraction)
      c->stats[2 + 5] += ms[5];            // EOM
      float lums = 0.f;
      CK(hipEventElapsedTime(&lums, c->ev_lu[0], c->ev_lu[1]));
      lane_store(c, lums);
      c->stats[9] += lums;                 // the batched LU kernel alone
      c->stats[10] += 1;                   // its launches
      c->stats[11] += c->lut.collect();    // the dominant GEMM kernel: milliseconds,
      c->stats[12] += c->lut.flops;        // algorithmic flops,
      c->stats[13] += c->lut.launches;     // launches
      c->stats[15] += nfailed;             // points left out (non-zero status)
      c->lut.flops = 0.0;
      c->stats[0] = std::max(c->stats[0], (double)iters);
      c->stats[1] = std::max(c->stats[1], maxrel);
      for (int b = 0; b < nb; ++b) {
        const int p = p0 + b;
        if (status) status[p] = st[b];
        if (st[b] != ADGA_OK && worst == ADGA_OK) worst = st[b];
        if (chid) memcpy(chid + 2 * (size_t)p * g.n4, &chih[(size_t)(2 * b) * g.n4], sizeof(z64) * g.n4);
        if (chim) memcpy(chim + 2 * (size_t)p * g.n4, &chih[(size_t)(2 * b + 1) * g.n4], sizeof(z64) * g.n4);
        if (bubnl) memcpy(bubnl + 2 * (size_t)p * g.n4, &bubh[(size_t)b * g.n4], sizeof(z64) * g.n4);
      }
    }
    if (worst == ADGA_E_SINGULAR) c->err = "singular BSE matrix (zero pivot in the LU)";
    else if (worst == ADGA_E_NOCONV) c->err = "iterative refinement did not reach the tolerance";
    else if (worst == ADGA_E_NONFINITE) c->err = "non-finite value in a BSE solve";
    return worst;
  });
}
}  // namespace

extern "C" {

int adga_run_points(adga_ctx* c, int npts, const int* iq, int precision, double* chid, double* chim,
                    double* bubnl, int* status) {
  return run_points_impl(c, npts, iq, precision, chid, chim, bubnl, status, nullptr);
}

int adga_run_points_tangent(adga_ctx* c, int npts, const int* iq, double* chid, double* chim, double* bubnl,
                            double* dchid, double* dchim, double* dbubnl, int* status) {
  if (!c) return ADGA_E_ARG;
  if (c->ndir == 0) return fail(c, ADGA_E_STATE, "adga_set_tangents must be called before adga_run_points_tangent");
  if (!c->t_omega || c->t_iwb != c->iwb || c->t_iwf != c->gw.iwf)
    return fail(c, ADGA_E_STATE, "adga_set_local_omega_tangent must follow the omega's adga_set_local_omega");
  if (c->vq) return fail(c, ADGA_E_ARG, "the derivatives are formed without V(q)");
  if (c->bubble_range != 0) return fail(c, ADGA_E_STATE, "the derivatives are formed with the bubble on the box");
  const TangentOut tan{dchid, dchim, dbubnl};
  return run_points_impl(c, npts, iq, ADGA_PREC_FP64, chid, chim, bubnl, status, &tan);
}

int adga_get_sigma(adga_ctx* c, double* sig_nl, double* sig_dmft, double* sig_hf) {
  if (!c) return ADGA_E_ARG;
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    eom_flush(c);
    adga::launch_fill_neg(g, g.nkp_eom, c->sig_nl, s);
    CK(hipGetLastError());
    if (sig_nl) CK(hipMemcpyAsync(sig_nl, c->sig_nl, sizeof(z64) * g.n2 * g.nf * g.nkp_eom, hipMemcpyDeviceToHost, s));
    if (sig_dmft) CK(hipMemcpyAsync(sig_dmft, c->sig_dmft, sizeof(z64) * g.n2 * g.nf, hipMemcpyDeviceToHost, s));
    if (sig_hf) CK(hipMemcpyAsync(sig_hf, c->sig_hf, sizeof(z64) * g.n2 * g.nkp_eom, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_get_giw(adga_ctx* c, double* giw) {
  if (!c || !giw) return ADGA_E_ARG;
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    const size_t ne = (size_t)2 * g.iwmax * g.n;
    DevBuf<> d(ne);
    const int rc = op_finish(c, adga::launch_op_giw(g, c->hk, c->siw, c->dc, d.p, c->stream), "get_giw");
    if (rc != ADGA_OK) return rc;
    // like get_giw, which overwrites the module's giw: the context's copy (the local bubble's input)
    // becomes the result, and the local data of the current omega must be formed again
    CK(hipMemcpyAsync(c->giw, d.p, sizeof(z64) * ne, hipMemcpyDeviceToDevice, c->stream));
    c->have_omega = false;
    CK(hipMemcpyAsync(giw, d.p, sizeof(z64) * ne, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return ADGA_OK;
  });
}

int adga_get_nfock(adga_ctx* c, double* n_fock) {
  if (!c || !n_fock) return ADGA_E_ARG;
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    const size_t ne = (size_t)g.nkp * g.n2;
    DevBuf<> d(ne);
    const int rc = op_finish(c, adga::launch_op_nfock(g, c->hk, c->siw, c->dc, d.p, c->stream), "get_nfock");
    if (rc != ADGA_OK) return rc;
    CK(hipMemcpyAsync(n_fock, d.p, sizeof(z64) * ne, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return ADGA_OK;
  });
}

// The context's own sigma_sum in `ds` (n2 * 2 iwfmax_small * nkp_eom): the accumulated Sigma_nl (the pending
// EOM convolved, the negative frequencies filled as adga_get_sigma does) plus siw on the orbital diagonal,
// as add_siw_dmft forms it (src/eom_module.f90:237-251, src/main.F90:869) - a device copy instead of the
// caller's array crossing PCIe.
void own_sigma_sum(adga_ctx* c, z64* ds) {
  const Geom& g = c->g;
  eom_flush(c);
  adga::launch_fill_neg(g, g.nkp_eom, c->sig_nl, c->stream);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(ds, c->sig_nl, sizeof(z64) * g.n2 * g.nf * g.nkp_eom, hipMemcpyDeviceToDevice, c->stream));
  adga::launch_add_siw(g, g.nkp_eom, c->siw, ds, c->stream);
  CK(hipGetLastError());
}

int adga_get_sigma_g_loc(adga_ctx* c, const double* sigma_sum, double* sigma_loc, double* gloc) {
  if (!c) return ADGA_E_ARG;
  const Geom& g = c->g;
  if (g.nkp_eom != g.nkp) {
    c->err = "get_sigma_g_loc: sigma_sum must cover the whole k grid (nkp_eom == nkp)";
    return ADGA_E_ARG;
  }
  return guarded(c, [&]() -> int {
    const size_t ns = (size_t)g.n2 * g.nf * g.nkp, nl = (size_t)g.n2 * g.nf, ng = (size_t)2 * g.iwmax * g.n2;
    DevBuf<> ds(ns), dl(nl), dg(ng), dw(adga::op_sigma_loc_scratch(g));
    if (sigma_sum)
      upload(ds.p, sigma_sum, ns, c->stream);
    else
      own_sigma_sum(c, ds.p);
    const int rc = op_finish(
        c, adga::launch_op_sigma_g_loc(g, c->hk, c->siw, c->dc, ds.p, dl.p, dg.p, dw.p, c->stream), "get_sigma_g_loc");
    if (rc != ADGA_OK) return rc;
    if (sigma_loc) CK(hipMemcpyAsync(sigma_loc, dl.p, sizeof(z64) * nl, hipMemcpyDeviceToHost, c->stream));
    if (gloc) CK(hipMemcpyAsync(gloc, dg.p, sizeof(z64) * ng, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return ADGA_OK;
  });
}

int adga_get_ndga(adga_ctx* c, const double* sigma_sum, double* n_dga_k, double* n_dga) {
  if (!c) return ADGA_E_ARG;
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    const size_t ns = (size_t)g.n2 * g.nf * g.nkp_eom, nk = (size_t)g.nkp_eom * g.n2;
    DevBuf<> ds(ns), dk(nk), dn((size_t)g.n);
    if (sigma_sum)
      upload(ds.p, sigma_sum, ns, c->stream);
    else
      own_sigma_sum(c, ds.p);
    const int rc = op_finish(
        c, adga::launch_op_ndga(g, c->hk, c->siw, c->dc, c->keom, ds.p, dk.p, dn.p, c->stream), "get_ndga");
    if (rc != ADGA_OK) return rc;
    if (n_dga_k) CK(hipMemcpyAsync(n_dga_k, dk.p, sizeof(z64) * nk, hipMemcpyDeviceToHost, c->stream));
    if (n_dga) CK(hipMemcpyAsync(n_dga, dn.p, sizeof(z64) * g.n, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return ADGA_OK;
  });
}

int adga_reset_sigma(adga_ctx* c) {
  if (!c) return ADGA_E_ARG;
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    CK(hipMemsetAsync(c->sig_nl, 0, sizeof(z64) * g.n2 * g.nf * g.nkp_eom, c->stream));
    CK(hipMemsetAsync(c->sig_dmft, 0, sizeof(z64) * g.n2 * g.nf, c->stream));
    CK(hipMemsetAsync(c->sig_hf, 0, sizeof(z64) * g.n2 * g.nkp_eom, c->stream));
    if (c->t_sig) CK(hipMemsetAsync(c->t_sig, 0, sizeof(z64) * g.n2 * g.nf * g.nkp_eom * c->ndir, c->stream));
    if (c->macc) CK(hipMemsetAsync(c->macc, 0, sizeof(z64) * (size_t)g.iwf * g.n4 * g.nkp, c->stream));
    if (c->t_macc)
      CK(hipMemsetAsync(c->t_macc, 0, sizeof(z64) * (size_t)g.iwf * g.n4 * g.nkp * c->ndir, c->stream));
    c->eom_pending = false;
    c->bub_valid = false;  // the next batch forms the bubble of its omega again
    c->t_bub_valid = false;
    CK(hipStreamSynchronize(c->stream));
    return ADGA_OK;
  });
}

int adga_get_stats(adga_ctx* c, double* out, int nout) {
  if (!c || !out || nout < 0) return ADGA_E_ARG;
  for (int i = 0; i < nout && i < ADGA_NSTATS; ++i) out[i] = c->stats[i];
  return ADGA_OK;
}

int adga_last_error(adga_ctx* c, char* buf, int len) {
  if (!c || !buf || len <= 0) return ADGA_E_ARG;
  snprintf(buf, (size_t)len, "%s", c->err.c_str());
  return ADGA_OK;
}

int adga_ctx_destroy(adga_ctx* c) {
  if (!c) return ADGA_OK;
  lane_store(c, INFINITY);  // a lane trial left pending by a failed batch counts as the slowest
  free_batch(c);
  free_fallback(c);
  dfree(c->hk); dfree(c->siw); dfree(c->giw); dfree(c->dc); dfree(c->u); dfree(c->ut); dfree(c->ndmft);
  dfree(c->gt); dfree(c->keom);
  dfree(c->xd); dfree(c->xm); dfree(c->cd); dfree(c->cm); dfree(c->gamd); dfree(c->gamm); dfree(c->opgd);
  dfree(c->opgm); dfree(c->winv); dfree(c->chi0w); dfree(c->chiloc_d); dfree(c->chiloc_m); dfree(c->bubloc);
  dfree(c->sig_nl); dfree(c->sig_dmft); dfree(c->sig_hf);
  dfree(c->vq); dfree(c->nfock); dfree(c->W); dfree(c->vqQ); dfree(c->qpos);
  for (z64*& p : c->stage) dfree(p);
  free_tangent(c);
  dfree(c->tw); dfree(c->ghat); dfree(c->bub_p); dfree(c->bub_all); dfree(c->macc); dfree(c->mwork);
  dfree(c->outer_p); dfree(c->outer_all); dfree(c->bubout);
  for (auto e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto e : c->ev_lu)
    if (e) (void)hipEventDestroy(e);
  for (int j = 0; j < adga::LU_MAX_LANES; ++j) {
    if (c->lanes.lag[j]) (void)hipEventDestroy(c->lanes.lag[j]);
    if (c->lanes.join[j]) (void)hipEventDestroy(c->lanes.join[j]);
    if (j > 0 && c->lanes.s[j]) (void)hipStreamDestroy(c->lanes.s[j]);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ADGA_OK;
}

int adga_inverse_matrix_z(int n, double* m, int* ierr) {
  if (n < 1 || !m || !ierr) return ADGA_E_ARG;
  *ierr = 0;
  if (n > ADGA_INVERSE_MAX_N) return ADGA_E_ARG;  // the batched LU's largest size (pivot table in LDS)
  return guarded(nullptr, [&]() -> int {
    const int W = 2 * n, ldm = (W + 15) / 16 * 16;
    std::vector<z64> h((size_t)n * ldm);
    const z64* src = reinterpret_cast<const z64*>(m);
    // row r of the row-major buffer = column r of the Fortran matrix, i.e. the factored
    // matrix is M^T; solving M^T Y = 1 gives Y = M^-T, whose rows are the columns of M^-1.
    for (int r = 0; r < n; ++r) {
      for (int cc = 0; cc < n; ++cc) h[(size_t)r * ldm + cc] = src[(size_t)r * n + cc];
      for (int cc = n; cc < ldm; ++cc) h[(size_t)r * ldm + cc] = adga::mk(cc - n == r ? 1.0 : 0.0, 0.0);
    }
    DevBuf<> dM(h.size()), dinv(adga::lu_inv_elems(n));
    DevBuf<int> dpiv((size_t)n), dinfo(1);
    CK(hipMemcpy(dM.p, h.data(), sizeof(z64) * h.size(), hipMemcpyHostToDevice));
    std::vector<int> ops = adga::lu_ops_factor(n, W);
    if (!adga::launch_lu<double>(1, n, W, ldm, ops, dM.p, dpiv.p, dinv.p, dinfo.p, nullptr)) return ADGA_E_ARG;
    CK(hipGetLastError());
    CK(hipMemcpy(h.data(), dM.p, sizeof(z64) * h.size(), hipMemcpyDeviceToHost));
    CK(hipMemcpy(ierr, dinfo.p, sizeof(int), hipMemcpyDeviceToHost));
    if (*ierr != 0) return ADGA_E_SINGULAR;
    z64* dst = reinterpret_cast<z64*>(m);
    for (int r = 0; r < n; ++r)
      for (int cc = 0; cc < n; ++cc) dst[(size_t)r * n + cc] = h[(size_t)r * ldm + n + cc];
    return ADGA_OK;
  });
}

}  // extern "C"
