// A8 (eta), A9 (chi contraction), A10 (dynamic EOM), A11 (local/static EOM) on gfx950.
#include "adga_kernels.h"

namespace adga {

// eta_m(s, x) = Y[m][x][s]: Y = eta^T solves A^T eta^T = B (main.F90:740-757), and the solutions' layout
// (one N x n2 row-major block per matrix) is eta's Fortran (n2, N) layout, so the kernels below read
// eta from Y itself.

// chi_ch(q) (non-local part) = beta^-2 gamma^w_ch . (chi0^q - chi0^w) + beta^-2 eta_ch . chi0^q, where
// (X . Y)(s, s2) = sum_{d, i3} X(s, d*n2 + i3) Y(i3, s2, d)  (susc_module.f90:34-56, main.F90:666-769);
// bubble_nl(q) = sum_d chi0^q / beta^2 - sum_d chi0^w / beta^2  (main.F90:654-663).
__global__ void k_chi(Geom g, int npts, const z64* __restrict__ chi0q, const z64* __restrict__ chi0w,
                      const z64* __restrict__ gam_d, const z64* __restrict__ gam_m, const z64* __restrict__ eta,
                      const z64* __restrict__ bubloc, z64* __restrict__ chi_out, z64* __restrict__ bub_out) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npts * 2 * g.n4) return;
  int e = (int)(idx % g.n4);
  int ch = (int)((idx / g.n4) % 2);
  int b = (int)(idx / (2 * g.n4));
  int s = e % g.n2, s2 = e / g.n2;
  const z64* gam = ch == 0 ? gam_d : gam_m;
  const z64* et = eta + (long)(2 * b + ch) * g.n2 * g.N;
  const z64* cq = chi0q + (long)b * g.nf * g.n4;
  z64 a1 = mk(0.0, 0.0), a2 = mk(0.0, 0.0), bs = mk(0.0, 0.0);
  for (int d = 0; d < g.nf; ++d) {
    for (int i3 = 0; i3 < g.n2; ++i3) {
      long x = (long)d * g.n2 + i3;
      z64 c = cq[(long)d * g.n4 + i3 + s2 * g.n2];
      z64 cnl = c;
      if (i3 == s2) cnl = csub(c, chi0w[s2 + d * g.n2]);
      a1 = cfma(gam[s + x * g.n2], cnl, a1);
      a2 = cfma(et[s + x * g.n2], c, a2);
    }
    if (ch == 0) bs = cadd(bs, cq[(long)d * g.n4 + s + s2 * g.n2]);
  }
  double ib2 = 1.0 / (g.beta * g.beta);
  z64 v = cadd(cscale(a1, ib2), cscale(a2, ib2));
  chi_out[((long)b * 2 + ch) * g.n4 + e] = v;
  if (ch == 0) bub_out[(long)b * g.n4 + e] = csub(cscale(bs, ib2), bubloc[e]);
}

// The same for n <= 3 with one workgroup per point: thread (k slice, channel, row s) accumulates row s of
// chi_ch over its slice of x = d*n2 + i3 (the n2 values of chi0^q at one x are read by every thread of the
// slice), the slices are summed through LDS, and threads 0 .. n4-1 form bubble_nl (summed over d in the
// same order as k_chi).  The chi sums run in another order than k_chi's: roundoff-level differences.
template <int N2>
__global__ void __launch_bounds__(256) k_chi_wg(Geom g, const z64* __restrict__ chi0q, const z64* __restrict__ chi0w,
                                                const z64* __restrict__ gam_d, const z64* __restrict__ gam_m,
                                                const z64* __restrict__ eta, const z64* __restrict__ bubloc,
                                                z64* __restrict__ chi_out, z64* __restrict__ bub_out) {
  constexpr int N4 = N2 * N2, KS = 256 / (2 * N2);
  __shared__ z64 part[KS][2][N4];
  const int b = blockIdx.x, t = threadIdx.x;
  const int ks = t / (2 * N2), ch = (t % (2 * N2)) / N2, s = t % N2;
  const int N = g.N;
  const z64* cq = chi0q + (long)b * g.nf * N4;
  if (ks < KS) {
    const z64* gam = ch == 0 ? gam_d : gam_m;
    const z64* et = eta + (long)(2 * b + ch) * N2 * N;
    z64 acc[N2];
#pragma unroll
    for (int s2 = 0; s2 < N2; ++s2) acc[s2] = mk(0.0, 0.0);
    const int x0 = (int)((long)N * ks / KS), x1 = (int)((long)N * (ks + 1) / KS);
#pragma unroll 2
    for (int x = x0; x < x1; ++x) {
      const int d = x / N2, i3 = x % N2;
      const z64 gv = gam[s + (long)x * N2], ev = et[s + (long)x * N2];
      const z64 w = chi0w[i3 + d * N2];  // chi0^w enters chi0^q - chi0^w where s2 == i3
      const z64* cx = cq + (long)d * N4 + i3;
#pragma unroll
      for (int s2 = 0; s2 < N2; ++s2) {
        const z64 c = cx[s2 * N2];
        acc[s2] = cfma(gv, s2 == i3 ? csub(c, w) : c, acc[s2]);
        acc[s2] = cfma(ev, c, acc[s2]);
      }
    }
#pragma unroll
    for (int s2 = 0; s2 < N2; ++s2) part[ks][ch][s + s2 * N2] = acc[s2];
  }
  __syncthreads();
  const double ib2 = 1.0 / (g.beta * g.beta);
  for (int e = t; e < 2 * N4; e += 256) {
    const int c2 = e / N4, f = e % N4;
    z64 v = mk(0.0, 0.0);
    for (int k = 0; k < KS; ++k) v = cadd(v, part[k][c2][f]);
    chi_out[((long)b * 2 + c2) * N4 + f] = cscale(v, ib2);
  }
  for (int e = t; e < N4; e += 256) {
    z64 bs = mk(0.0, 0.0);
#pragma unroll 8
    for (int d = 0; d < g.nf; ++d) bs = cadd(bs, cq[(long)d * N4 + e]);
    bub_out[(long)b * N4 + e] = csub(cscale(bs, ib2), bubloc[e]);
  }
}

// m_tot = -(U + V - U~/2) eta_d /(beta nqp) + 1.5 U~ eta_m /(beta nqp) + U gamma^q_d /(beta nqp)
//        - V gamma^w_d /(beta nqp)   (eom_module.f90:183-207; the V terms only with V(q): vq non-null).
// Output (n2, N) Fortran per point.
__global__ void k_mtot(Geom g, int npts, const z64* __restrict__ u, const z64* __restrict__ ut,
                       const z64* __restrict__ eta, const z64* __restrict__ gamq, z64* __restrict__ mt,
                       const int* __restrict__ qpos, const z64* __restrict__ vq, const z64* __restrict__ gamd,
                       int accumulate) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long per = (long)g.n2 * g.N;
  if (idx >= per * npts) return;
  int b = (int)(idx / per);
  long e = idx % per;
  int s = (int)(e % g.n2);
  long x = e / g.n2;
  const z64* ed = eta + (long)(2 * b) * per;
  const z64* em = eta + (long)(2 * b + 1) * per;
  const z64* gq = gamq + (long)(2 * b) * per;  // residual layout: point b at slot 2b
  const z64* v = vq ? vq + (long)qpos[b] * g.n4 : nullptr;
  z64 t1 = mk(0.0, 0.0), t2 = mk(0.0, 0.0), t3 = mk(0.0, 0.0), t4 = mk(0.0, 0.0);
  for (int a = 0; a < g.n2; ++a) {
    z64 uu = u[s + a * g.n2], uw = ut[s + a * g.n2];
    z64 uwork = mk(uu.re - 0.5 * uw.re, uu.im - 0.5 * uw.im);
    if (v) {
      const z64 va = v[s + a * g.n2];
      uwork = cadd(va, uwork);
      t4 = cfma(va, gamd[a + x * g.n2], t4);
    }
    t1 = cfma(uwork, ed[a + x * g.n2], t1);
    t2 = cfma(uw, em[a + x * g.n2], t2);
    t3 = cfma(uu, gq[a + x * g.n2], t3);
  }
  double a1 = -1.0 / g.beta / g.nqp, a2 = 1.5 / g.beta / g.nqp, a3 = 1.0 / g.beta / g.nqp;
  z64 m = cscale(t1, a1);
  m = cadd(cscale(t2, a2), m);
  m = cadd(cscale(t3, a3), m);
  if (v) m = cadd(cscale(t4, a1), m);
  mt[idx] = accumulate ? cadd(mt[idx], m) : m;
}

// Sigma_nl(i, l, nu, k) += sum_b sum_{j,k'} m_b(i*n + j, n2*(iwf + nu) + l*n + k') G(k - q_b, nu - omega)_{jk'}
// for nu = 0 .. iwf-1 (eom_module.f90:211-227).  One thread per (i, l, nu, k); the batch is summed
// in a fixed order, so the accumulation is deterministic.
__global__ void __launch_bounds__(256) k_eom(Geom g, int iwb, int npts, const int* __restrict__ qidx,
                                             const int* __restrict__ keom, const z64* __restrict__ mt,
                                             const z64* __restrict__ gt, z64* __restrict__ sigma) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int half = g.iwf;
  if (idx >= (long)g.n2 * half * g.nkp_eom) return;
  int il = (int)(idx % g.n2);
  int nu = (int)((idx / g.n2) % half);
  int kk = (int)(idx / ((long)g.n2 * half));
  int i = il % g.n, l = il / g.n;
  int k = keom[kk] - 1;
  const z64* gw = gt + (long)(nu - iwb - g.glo) * g.nkp * g.n2;
  long col0 = (long)g.n2 * (g.iwf + nu) + l * g.n;
  long per = (long)g.n2 * g.N;
  z64 acc = mk(0.0, 0.0);
  for (int b = 0; b < npts; ++b) {
    int kq = k_minus_q(k, qidx[b], g.nkx, g.nky, g.nkz);
    const z64* gr = gw + (long)kq * g.n2;
    const z64* m = mt + (long)b * per;
    for (int j = 0; j < g.n; ++j)
      for (int kp = 0; kp < g.n; ++kp) acc = cfma(m[(i * g.n + j) + (col0 + kp) * g.n2], gr[j * g.n + kp], acc);
  }
  long o = i + (long)l * g.n + (long)(g.iws + nu) * g.n2 + (long)kk * g.n2 * (2 * g.iws);
  sigma[o] = cadd(sigma[o], acc);
}

// Sigma(i, l, -nu-1) = conj(Sigma(l, i, nu)) for nu >= 0 (eom_module.f90:157, 228); nk slices.
__global__ void k_fill_neg(Geom g, int nk, z64* __restrict__ sigma) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)g.n2 * g.iws * nk) return;
  int il = (int)(idx % g.n2);
  int nu = (int)((idx / g.n2) % g.iws);
  long kk = idx / ((long)g.n2 * g.iws);
  int i = il % g.n, l = il / g.n;
  long base = kk * g.n2 * (2 * g.iws);
  z64 v = sigma[base + l + (long)i * g.n + (long)(g.iws + nu) * g.n2];
  sigma[base + i + (long)l * g.n + (long)(g.iws - nu - 1) * g.n2] = cconj(v);
}

// Local EOM with the DMFT g: Sigma_dmft(i, l, nu) += sum_k m(i*n + k, n2*(iwf+nu) + l*n + k) g_k(nu - omega),
// m = -(1/beta) U gamma^w_d (eom_module.f90:137-158, default branch).
__global__ void k_eom_dmft(Geom g, int iwb, const z64* __restrict__ u, const z64* __restrict__ gam_d,
                           const z64* __restrict__ giw, z64* __restrict__ sdmft) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.n2 * g.iwf) return;
  int il = idx % g.n2, nu = idx / g.n2;
  int i = il % g.n, l = il / g.n;
  long col0 = (long)g.n2 * (g.iwf + nu) + l * g.n;
  z64 acc = mk(0.0, 0.0);
  for (int k = 0; k < g.n; ++k) {
    int a = i * g.n + k;
    z64 t = mk(0.0, 0.0);
    for (int ap = 0; ap < g.n2; ++ap) t = cfma(u[a + ap * g.n2], gam_d[ap + (col0 + k) * g.n2], t);
    z64 m = cscale(t, -1.0 / g.beta);
    z64 gl = giw[(nu - iwb + g.iwmax) + (long)k * 2 * g.iwmax];
    acc = cfma(m, gl, acc);
  }
  long o = i + (long)l * g.n + (long)(g.iws + nu) * g.n2;
  sdmft[o] = cadd(sdmft[o], acc);
}

// Local Hartree term sum_j (2 u({ii},{jj}) - u({ji},{ji})) n_j added to every frequency
// (eom_module.f90:58-70).
__global__ void k_eom_static(Geom g, const z64* __restrict__ u, const z64* __restrict__ ndmft,
                             z64* __restrict__ sdmft) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.n * 2 * g.iws) return;
  int i = idx % g.n, nu = idx / g.n;
  z64 acc = mk(0.0, 0.0);
  for (int j = 0; j < g.n; ++j) {
    z64 a = u[(i * g.n + i) + (long)(j * g.n + j) * g.n2];
    z64 b = u[(i * g.n + j) + (long)(i * g.n + j) * g.n2];
    acc = cfma(mk(2.0 * a.re - b.re, 2.0 * a.im - b.im), ndmft[j], acc);
  }
  long o = i + (long)i * g.n + (long)nu * g.n2;
  sdmft[o] = cadd(sdmft[o], acc);
}

void launch_chi(const Geom& g, int npts, const z64* eta, const z64* chi0q, const z64* chi0w, const z64* gamd,
                const z64* gamm, const z64* bubloc, z64* chi_out, z64* bub_out, hipStream_t s) {
  if (g.n <= 3) {
#define ADGA_CHI_WG(NN)                                                                                          \
  if (g.n == NN) {                                                                                              \
    hipLaunchKernelGGL(k_chi_wg<NN * NN>, dim3((unsigned)npts), dim3(256), 0, s, g, chi0q, chi0w, gamd, gamm, eta, \
                       bubloc, chi_out, bub_out);                                                               \
    return;                                                                                                     \
  }
    ADGA_CHI_WG(1) ADGA_CHI_WG(2) ADGA_CHI_WG(3)
#undef ADGA_CHI_WG
  }
  long t2 = (long)npts * 2 * g.n4;
  hipLaunchKernelGGL(k_chi, dim3((unsigned)((t2 + 127) / 128)), dim3(128), 0, s, g, npts, chi0q, chi0w, gamd, gamm,
                     eta, bubloc, chi_out, bub_out);
}

// k_mtot for n <= 3 with one thread per (point b, x) forming the n2 values s of its row: eta_d, eta_m and
// gamma^q_d at x (and gamma^w_d with V(q)) are read once per thread instead of once per output; the same
// sums in the same order as k_mtot.
template <int N2, bool VQ>
__global__ void __launch_bounds__(256) k_mtot_t(Geom g, int npts, const z64* __restrict__ u,
                                                const z64* __restrict__ ut, const z64* __restrict__ eta,
                                                const z64* __restrict__ gamq, z64* __restrict__ mt,
                                                const int* __restrict__ qpos, const z64* __restrict__ vq,
                                                const z64* __restrict__ gamd, int accumulate) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npts * g.N) return;
  const int b = (int)(idx / g.N), x = (int)(idx - (long)b * g.N);
  const long per = (long)N2 * g.N;
  const z64* ed = eta + (long)(2 * b) * per + (long)x * N2;
  const z64* em = eta + (long)(2 * b + 1) * per + (long)x * N2;
  const z64* gq = gamq + (long)(2 * b) * per + (long)x * N2;  // residual layout: point b at slot 2b
  z64 e_d[N2], e_m[N2], g_q[N2], g_d[VQ ? N2 : 1];
#pragma unroll
  for (int a = 0; a < N2; ++a) {
    e_d[a] = ed[a];
    e_m[a] = em[a];
    g_q[a] = gq[a];
    if constexpr (VQ) g_d[a] = gamd[a + (long)x * N2];
  }
  const z64* v = VQ ? vq + (long)qpos[b] * (N2 * N2) : nullptr;
  const double a1 = -1.0 / g.beta / g.nqp, a2 = 1.5 / g.beta / g.nqp, a3 = 1.0 / g.beta / g.nqp;
  z64* out = mt + (long)b * per + (long)x * N2;
#pragma unroll
  for (int s = 0; s < N2; ++s) {
    z64 t1 = mk(0.0, 0.0), t2 = mk(0.0, 0.0), t3 = mk(0.0, 0.0), t4 = mk(0.0, 0.0);
#pragma unroll
    for (int a = 0; a < N2; ++a) {
      const z64 uu = u[s + a * N2], uw = ut[s + a * N2];
      z64 uwork = mk(uu.re - 0.5 * uw.re, uu.im - 0.5 * uw.im);
      if constexpr (VQ) {
        const z64 va = v[s + a * N2];
        uwork = cadd(va, uwork);
        t4 = cfma(va, g_d[a], t4);
      }
      t1 = cfma(uwork, e_d[a], t1);
      t2 = cfma(uw, e_m[a], t2);
      t3 = cfma(uu, g_q[a], t3);
    }
    z64 m = cscale(t1, a1);
    m = cadd(cscale(t2, a2), m);
    m = cadd(cscale(t3, a3), m);
    if constexpr (VQ) m = cadd(cscale(t4, a1), m);
    out[s] = accumulate ? cadd(out[s], m) : m;
  }
}

void launch_mtot(const Geom& g, int npts, const z64* u, const z64* ut, const z64* eta, const z64* gamq, z64* mt,
                 hipStream_t s, const int* qpos, const z64* vq, const z64* gamd, bool accumulate) {
  const long tx = (long)npts * g.N;
#define ADGA_MT(NN)                                                                                                 \
  if (g.n == NN) {                                                                                                 \
    if (vq)                                                                                                        \
      hipLaunchKernelGGL((k_mtot_t<NN * NN, true>), dim3((unsigned)((tx + 255) / 256)), dim3(256), 0, s, g, npts,  \
                         u, ut, eta, gamq, mt, qpos, vq, gamd, accumulate ? 1 : 0);                                 \
    else                                                                                                           \
      hipLaunchKernelGGL((k_mtot_t<NN * NN, false>), dim3((unsigned)((tx + 255) / 256)), dim3(256), 0, s, g, npts, \
                         u, ut, eta, gamq, mt, qpos, vq, gamd, accumulate ? 1 : 0);                                 \
    return;                                                                                                        \
  }
  ADGA_MT(1) ADGA_MT(2) ADGA_MT(3)
#undef ADGA_MT
  long tot = (long)npts * g.n2 * g.N;
  hipLaunchKernelGGL(k_mtot, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, u, ut, eta, gamq, mt,
                     qpos, vq, gamd, accumulate ? 1 : 0);
}

void launch_eom_direct(const Geom& g, int iwb, int npts, const int* qidx, const int* keom, const z64* mt,
                       const z64* gt, z64* sigma, hipStream_t s) {
  long te = (long)g.n2 * g.iwf * g.nkp_eom;
  hipLaunchKernelGGL(k_eom, dim3((unsigned)((te + 255) / 256)), dim3(256), 0, s, g, iwb, npts, qidx, keom, mt, gt,
                     sigma);
}

void launch_eom_static_local(const Geom& g, const z64* u, const z64* ndmft, z64* sdmft, hipStream_t s) {
  hipLaunchKernelGGL(k_eom_static, dim3((g.n * 2 * g.iws + 63) / 64), dim3(64), 0, s, g, u, ndmft, sdmft);
}

void launch_eom_local(const Geom& g, int iwb, bool do_static, const z64* u, const z64* gamd, const z64* giw,
                      const z64* ndmft, z64* sdmft, hipStream_t s) {
  if (do_static)
    hipLaunchKernelGGL(k_eom_static, dim3((g.n * 2 * g.iws + 63) / 64), dim3(64), 0, s, g, u, ndmft, sdmft);
  hipLaunchKernelGGL(k_eom_dmft, dim3((g.n2 * g.iwf + 63) / 64), dim3(64), 0, s, g, iwb, u, gamd, giw, sdmft);
  hipLaunchKernelGGL(k_fill_neg, dim3((g.n2 * g.iws + 63) / 64), dim3(64), 0, s, g, 1, sdmft);
}

__global__ void k_zero_points(int npts, long len, const int* __restrict__ keep, z64* __restrict__ a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npts * len) return;
  if (!keep[idx / len]) a[idx] = mk(0.0, 0.0);
}

void launch_zero_points(int npts, long len, const int* keep, z64* a, hipStream_t s) {
  const long tot = (long)npts * len;
  if (tot > 0)
    hipLaunchKernelGGL(k_zero_points, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, npts, len, keep, a);
}

__global__ void k_gather(int k, long len, const int* __restrict__ sel, const z64* __restrict__ src,
                         z64* __restrict__ dst, int scatter) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)k * len) return;
  const long i = idx / len, e = idx % len;
  const long far = (long)sel[i] * len + e;
  if (scatter) dst[far] = src[idx];
  else dst[idx] = src[far];
}

void launch_gather(int k, long len, const int* sel, const z64* src, z64* dst, bool scatter, hipStream_t s) {
  const long tot = (long)k * len;
  if (tot > 0)
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, k, len, sel, src, dst,
                       scatter ? 1 : 0);
}

void launch_fill_neg(const Geom& g, int nk, z64* sigma, hipStream_t s) {
  long tot = (long)g.n2 * g.iws * nk;
  hipLaunchKernelGGL(k_fill_neg, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, nk, sigma);
}

}  // namespace adga
