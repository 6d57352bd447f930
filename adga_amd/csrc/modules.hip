// Kernels behind the module-level entry points (one call per call of a reference module procedure):
// the local bubble's diagonal for every nu of the box (get_chi0_loc, get_chi0_loc_inv), one lattice
// Green's function (get_gkiw, get_gkiw_dga), the chi contraction on caller data (calc_chi_qw), n_dmft
// (get_ndmft) and the DMFT self-energy added to the DGA one (add_siw_dmft).
#include "adga_kernels.h"
#include "small_inv.h"

namespace adga {

// diag(i1, d) with i1 = i*n + j: chi0^w = -beta g_i(nu) g_j(nu - omega) and its reciprocal
// (one_particle_quant_module.f90:192-225), nu = nu_{d - iwf} over the box.
__global__ void k_chi0_loc_diag(Geom g, int iwb, const z64* __restrict__ giw, z64* __restrict__ w,
                                z64* __restrict__ winv) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.n2 * g.nf) return;
  const int i1 = idx % g.n2, d = idx / g.n2;
  const int i = i1 / g.n, j = i1 % g.n;
  const int iw = d - g.iwf;
  const z64 gi = giw[(iw + g.iwmax) + (long)i * 2 * g.iwmax];
  const z64 gj = giw[(iw - iwb + g.iwmax) + (long)j * 2 * g.iwmax];
  const z64 p = cmul(gi, gj);
  if (w) w[idx] = cscale(p, -g.beta);
  if (winv) winv[idx] = cscale(crecip(p), -1.0 / g.beta);
}

void launch_chi0_loc_diag(const Geom& g, int iwb, const z64* giw, z64* w, z64* winv, hipStream_t s) {
  const int tot = g.n2 * g.nf;
  hipLaunchKernelGGL(k_chi0_loc_diag, dim3((tot + 63) / 64), dim3(64), 0, s, g, iwb, giw, w, winv);
}

// G(k, nu_iw) = [(i nu + mu) - H(k) - dc - Sigma]^-1 with Sigma = siw(iw) on the diagonal (skiw null,
// get_gkiw 81-104) or the caller's n x n skiw (get_gkiw_dga 106-127); out(n, n) in Fortran order.
template <int NN>
__global__ void k_gkiw(Geom g, int k, int iw, const z64* __restrict__ hk, const z64* __restrict__ siw,
                       const z64* __restrict__ dc, const z64* __restrict__ skiw, z64* __restrict__ out) {
  if (threadIdx.x != 0) return;
  const double nu = M_PI * (2 * iw + 1) / g.beta;
  z64 a[NN][NN];
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      z64 h = hk[(long)k * NN * NN + i + j * NN];
      if (skiw) h = cadd(h, skiw[i + j * NN]);
      a[i][j] = mk(-h.re, -h.im);
    }
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const z64 d = dc[2 * i];
    z64 v = mk(g.mu - d.re, nu - d.im);
    if (!skiw) v = csub(v, siw[(iw + g.iwmax) + (long)i * 2 * g.iwmax]);
    a[i][i] = cadd(a[i][i], v);
  }
  inv_small<NN>(a);
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int j = 0; j < NN; ++j) out[i + j * NN] = a[i][j];
}

bool launch_gkiw(const Geom& g, int k, int iw, const z64* hk, const z64* siw, const z64* dc, const z64* skiw,
                 z64* out, hipStream_t s) {
  switch (g.n) {
#define ADGA_GK(NN) \
  case NN: hipLaunchKernelGGL(k_gkiw<NN>, dim3(1), dim3(64), 0, s, g, k, iw, hk, siw, dc, skiw, out); return true;
    ADGA_GK(1) ADGA_GK(2) ADGA_GK(3) ADGA_GK(4) ADGA_GK(5) ADGA_GK(6) ADGA_GK(7) ADGA_GK(8)
#undef ADGA_GK
    default: return false;
  }
}

// chi(s, s2) += beta^-2 sum_{d, i3} X(s, d*n2 + i3) Y(i3, s2, d)  (susc_module.f90:34-56); X (n2, N),
// Y (n2, n2, nf), chi (n2, n2), all Fortran order.  One thread per output.
__global__ void k_chi_qw(Geom g, const z64* __restrict__ X, const z64* __restrict__ Y, z64* __restrict__ chi) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n4) return;
  const int s = e % g.n2, s2 = e / g.n2;
  z64 acc = mk(0.0, 0.0);
  for (int d = 0; d < g.nf; ++d)
    for (int i3 = 0; i3 < g.n2; ++i3)
      acc = cfma(X[s + (long)(d * g.n2 + i3) * g.n2], Y[i3 + (long)s2 * g.n2 + (long)d * g.n4], acc);
  chi[e] = cadd(chi[e], cscale(acc, 1.0 / (g.beta * g.beta)));
}

void launch_chi_qw(const Geom& g, const z64* X, const z64* Y, z64* chi, hipStream_t s) {
  hipLaunchKernelGGL(k_chi_qw, dim3((g.n4 + 63) / 64), dim3(64), 0, s, g, X, Y, chi);
}

// n_dmft(i) = 2/beta sum_{iw = 0}^{iwmax-1} Re giw(iw, i) + 1/2 (one_particle_quant_module.f90:295-313),
// one wave per orbital.
__global__ void __launch_bounds__(64) k_ndmft(Geom g, const z64* __restrict__ giw, z64* __restrict__ nd) {
  const int i = blockIdx.x, lane = threadIdx.x;
  double acc = 0.0;
  for (int iw = lane; iw < g.iwmax; iw += 64) acc += giw[(iw + g.iwmax) + (long)i * 2 * g.iwmax].re;
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if (lane == 0) nd[i] = mk(2.0 * acc / g.beta + 0.5, 0.0);
}

void launch_ndmft(const Geom& g, const z64* giw, z64* nd, hipStream_t s) {
  hipLaunchKernelGGL(k_ndmft, dim3(g.n), dim3(64), 0, s, g, giw, nd);
}

// sigma_sum(b, b, nu, k) += siw(nu, b) for nu in the box (eom_module.f90:237-251); sigma_sum (n, n, nf, nk).
__global__ void k_add_siw(Geom g, int nk, const z64* __restrict__ siw, z64* __restrict__ ss) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int nfs = 2 * g.iws;
  if (idx >= (long)g.n * nfs * nk) return;
  const int b = (int)(idx % g.n);
  const int d = (int)((idx / g.n) % nfs);
  const long k = idx / ((long)g.n * nfs);
  z64* o = ss + b + (long)b * g.n + (long)d * g.n2 + k * g.n2 * nfs;
  *o = cadd(*o, siw[(d - g.iws + g.iwmax) + (long)b * 2 * g.iwmax]);
}

void launch_add_siw(const Geom& g, int nk, const z64* siw, z64* ss, hipStream_t s) {
  const long tot = (long)g.n * 2 * g.iws * nk;
  hipLaunchKernelGGL(k_add_siw, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, nk, siw, ss);
}

}  // namespace adga
