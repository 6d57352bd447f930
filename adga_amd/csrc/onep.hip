// One-particle routines around the per-(q,omega) loop (one_particle_quant_module.f90) on gfx950:
// get_giw (40-79), get_nfock (337-361), get_sigma_g_loc (124-180) and get_ndga (363-393).  Each is
// a sum of small inverses G(k, nu) = [(i nu + mu) - H(k) - dc - Sigma(k, nu)]^-1 over the k grid or
// over the frequencies nu_0 .. nu_{iwmax-1}: one thread per (k, nu) inverts in registers, one
// workgroup per fixed index reduces (wave shuffles, then the four waves through LDS).
//
// Layouts are the reference's Fortran arrays (first index fastest): hk(n,n,nkp),
// siw(-iwmax:iwmax-1, n), dc(2, n), sigma_sum(n, n, -iwfmax_small:iwfmax_small-1, nk),
// giw(-iwmax:iwmax-1, n), gloc(-iwmax:iwmax-1, n, n), n_fock / n_dga_k(nk, n, n), n_dga(n).
#include "adga_kernels.h"
#include "small_inv.h"

namespace adga {

namespace {

constexpr int OP_NT = 256;

// G^-1(k, nu): -H(k) - Sigma off the diagonal, (i nu + mu) - H(k)_ii - dc(1,i) - Sigma_ii on it.
// Sigma is the DMFT siw (diagonal, frequency iw) when sfull is null, else the full matrix
// sfull(a, b) = sfull[a + b*n] (get_gkiw 81-104, get_gkiw_dga 107-121, get_sigma_g_loc 153-160).
template <int NN>
__device__ __forceinline__ void ginv(const Geom& g, const z64* __restrict__ hk, int k, int iw,
                                     const z64* __restrict__ siw, const z64* __restrict__ dc,
                                     const z64* __restrict__ sfull, z64 (&a)[NN][NN]) {
  const double nu = M_PI * (2 * iw + 1) / g.beta;
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      const z64 h = hk[(long)k * NN * NN + i + j * NN];
      z64 v = mk(-h.re, -h.im);
      if (sfull) v = csub(v, sfull[i + j * NN]);
      a[i][j] = v;
    }
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const z64 h = hk[(long)k * NN * NN + i + i * NN];
    const z64 s = sfull ? sfull[i + i * NN] : siw[(iw + g.iwmax) + (long)i * 2 * g.iwmax];
    const z64 d = dc[2 * i];
    a[i][i] = mk(g.mu - h.re - d.re - s.re, nu - h.im - d.im - s.im);
  }
}

// Sum of v over the workgroup, returned in thread 0 (red: one slot per wave).
__device__ __forceinline__ z64 wg_sum(z64 v, z64* red) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v.re += __shfl_xor(v.re, off);
    v.im += __shfl_xor(v.im, off);
  }
  __syncthreads();  // red is free (the previous sum has been read)
  if (lane == 0) red[wave] = v;
  __syncthreads();
  z64 s = mk(0.0, 0.0);
  if (t == 0)
    for (int w = 0; w < OP_NT / 64; ++w) s = cadd(s, red[w]);
  return s;
}

// k sums at fixed nu_iw (blockIdx.x = iw in [0, iwmax)), threads over k.
//   FULL = false: giw(iw, i) = (1/nkp) sum_k G_ii, giw(-iw-1, i) = conj (get_giw).
//   FULL = true:  gloc(iw, a, b) = (1/nkp) sum_k G_ab with Sigma = sigma_sum(:, :, iw, k) for
//                 iw < iwfmax_small and siw outside; gloc(-iw-1) = gloc(iw)^H (get_sigma_g_loc).
template <int NN, bool FULL>
__global__ void __launch_bounds__(OP_NT) k_op_ksum(Geom g, const z64* __restrict__ hk, const z64* __restrict__ siw,
                                                   const z64* __restrict__ dc, const z64* __restrict__ ssum,
                                                   z64* __restrict__ out) {
  __shared__ z64 red[OP_NT / 64];
  const int iw = blockIdx.x, t = threadIdx.x;
  constexpr int NE = FULL ? NN * NN : NN;
  z64 acc[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) acc[e] = mk(0.0, 0.0);
  const bool box = FULL && iw < g.iwf;
  for (int k = t; k < g.nkp; k += OP_NT) {
    const z64* sf = box ? ssum + ((long)k * g.nf + (iw + g.iwf)) * NN * NN : nullptr;
    z64 a[NN][NN];
    ginv<NN>(g, hk, k, iw, siw, dc, sf, a);
    inv_small<NN>(a);
    if constexpr (FULL) {
#pragma unroll
      for (int i = 0; i < NN; ++i)
#pragma unroll
        for (int j = 0; j < NN; ++j) acc[i + j * NN] = cadd(acc[i + j * NN], a[i][j]);
    } else {
#pragma unroll
      for (int i = 0; i < NN; ++i) acc[i] = cadd(acc[i], a[i][i]);
    }
  }
  const long w2 = 2L * g.iwmax;
  const double f = (double)g.nkp;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const z64 s = wg_sum(acc[e], red);
    if (t == 0) {
      const z64 v = mk(s.re / f, s.im / f);
      if constexpr (FULL) {
        const int i = e % NN, j = e / NN;
        out[(iw + g.iwmax) + i * w2 + j * w2 * NN] = v;
        out[(-iw - 1 + g.iwmax) + j * w2 + i * w2 * NN] = mk(v.re, -v.im);
      } else {
        out[(iw + g.iwmax) + e * w2] = v;
        out[(-iw - 1 + g.iwmax) + e * w2] = mk(v.re, -v.im);
      }
    }
  }
}

// frequency sums at fixed k (blockIdx.x = index into the k list), threads over nu_0 .. nu_{iwmax-1}:
// out(kk, a, b) = 2/beta sum_nu Re G_ab + delta_ab / 2 (complex, zero imaginary part).
//   DGA = false: k = kk on the full grid, Sigma = siw (get_nfock).
//   DGA = true:  k = k_data_eom(kk), Sigma = sigma_sum(:, :, iw, kk) for iw < iwfmax_small, siw
//                outside (get_ndga).
template <int NN, bool DGA>
__global__ void __launch_bounds__(OP_NT) k_op_wsum(Geom g, int nk, const z64* __restrict__ hk, const z64* __restrict__ siw,
                                                   const z64* __restrict__ dc, const int* __restrict__ keom,
                                                   const z64* __restrict__ ssum, z64* __restrict__ out) {
  __shared__ z64 red[OP_NT / 64];
  const int kk = blockIdx.x, t = threadIdx.x;
  const int k = DGA ? keom[kk] - 1 : kk;
  double acc[NN * NN];
#pragma unroll
  for (int e = 0; e < NN * NN; ++e) acc[e] = 0.0;
  for (int iw = t; iw < g.iwmax; iw += OP_NT) {
    const z64* sf = (DGA && iw < g.iwf) ? ssum + ((long)kk * g.nf + (iw + g.iwf)) * NN * NN : nullptr;
    z64 a[NN][NN];
    ginv<NN>(g, hk, k, iw, siw, dc, sf, a);
    inv_small<NN>(a);
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
      for (int j = 0; j < NN; ++j) acc[i + j * NN] += a[i][j].re;
  }
#pragma unroll
  for (int e = 0; e < NN * NN; ++e) {
    const z64 s = wg_sum(mk(acc[e], 0.0), red);
    if (t == 0) {
      const int i = e % NN, j = e / NN;
      double v = 2.0 * s.re / g.beta;
      if (i == j) v += 0.5;
      out[kk + (long)i * nk + (long)j * nk * NN] = mk(v, 0.0);
    }
  }
}

// n_dga(i) = (1/nk) sum_kk n_dga_k(kk, i, i) (get_ndga 383-389); one workgroup per orbital.
__global__ void __launch_bounds__(OP_NT) k_op_ndga_mean(int n, int nk, const z64* __restrict__ ndk, z64* __restrict__ nd) {
  __shared__ z64 red[OP_NT / 64];
  const int i = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int kk = t; kk < nk; kk += OP_NT) s += ndk[kk + (long)i * nk + (long)i * nk * n].re;
  const z64 r = wg_sum(mk(s, 0.0), red);
  if (t == 0) nd[i] = mk(r.re / nk, 0.0);
}

// sigma_loc(a, b, iw) = (1/nkp) sum_k sigma_sum(a, b, iw, k) over the whole box (get_sigma_g_loc 145-150),
// in two passes: partial sums over SL_KCH chunks of k (a wave reads 64 consecutive elements of one
// k, four waves take every fourth k of the chunk), then the sum of the partials.
constexpr int SL_KCH = 32;

__global__ void __launch_bounds__(OP_NT) k_op_sigma_loc_part(Geom g, const z64* __restrict__ ssum,
                                                             z64* __restrict__ part) {
  __shared__ z64 red[OP_NT / 64][64];
  const int t = threadIdx.x, el = t & 63, ks = t >> 6;
  const long ne = (long)g.n2 * g.nf;
  const long e = (long)blockIdx.x * 64 + el;
  const int kper = (g.nkp + SL_KCH - 1) / SL_KCH;
  const int k0 = blockIdx.y * kper, k1 = min(g.nkp, k0 + kper);
  z64 s = mk(0.0, 0.0);
  if (e < ne)
    for (int k = k0 + ks; k < k1; k += OP_NT / 64) s = cadd(s, ssum[(long)k * ne + e]);
  red[ks][el] = s;
  __syncthreads();
  if (ks == 0 && e < ne) {
    z64 r = red[0][el];
    for (int w = 1; w < OP_NT / 64; ++w) r = cadd(r, red[w][el]);
    part[(long)blockIdx.y * ne + e] = r;
  }
}

__global__ void k_op_sigma_loc_sum(Geom g, const z64* __restrict__ part, z64* __restrict__ sloc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ne = (long)g.n2 * g.nf;
  if (e >= ne) return;
  z64 s = mk(0.0, 0.0);
  for (int c = 0; c < SL_KCH; ++c) s = cadd(s, part[(long)c * ne + e]);
  sloc[e] = mk(s.re / g.nkp, s.im / g.nkp);
}

}  // namespace

#define ADGA_OP_SWITCH(F)                       \
  switch (g.n) {                               \
    case 1: F(1); break;                       \
    case 2: F(2); break;                       \
    case 3: F(3); break;                       \
    case 4: F(4); break;                       \
    case 5: F(5); break;                       \
    case 6: F(6); break;                       \
    case 7: F(7); break;                       \
    case 8: F(8); break;                       \
    default: return false;                     \
  }

bool launch_op_giw(const Geom& g, const z64* hk, const z64* siw, const z64* dc, z64* giw, hipStream_t s) {
#define F(NN) hipLaunchKernelGGL((k_op_ksum<NN, false>), dim3(g.iwmax), dim3(OP_NT), 0, s, g, hk, siw, dc, nullptr, giw)
  ADGA_OP_SWITCH(F)
#undef F
  return true;
}

bool launch_op_nfock(const Geom& g, const z64* hk, const z64* siw, const z64* dc, z64* nfock, hipStream_t s) {
#define F(NN)                                                                                                 \
  hipLaunchKernelGGL((k_op_wsum<NN, false>), dim3(g.nkp), dim3(OP_NT), 0, s, g, g.nkp, hk, siw, dc, nullptr, \
                     nullptr, nfock)
  ADGA_OP_SWITCH(F)
#undef F
  return true;
}

size_t op_sigma_loc_scratch(const Geom& g) { return (size_t)SL_KCH * g.n2 * g.nf; }

bool launch_op_sigma_g_loc(const Geom& g, const z64* hk, const z64* siw, const z64* dc, const z64* ssum, z64* sloc,
                           z64* gloc, z64* scratch, hipStream_t s) {
  const long ne = (long)g.n2 * g.nf;
  hipLaunchKernelGGL(k_op_sigma_loc_part, dim3((unsigned)((ne + 63) / 64), SL_KCH), dim3(OP_NT), 0, s, g, ssum,
                     scratch);
  hipLaunchKernelGGL(k_op_sigma_loc_sum, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, g, scratch, sloc);
#define F(NN) hipLaunchKernelGGL((k_op_ksum<NN, true>), dim3(g.iwmax), dim3(OP_NT), 0, s, g, hk, siw, dc, ssum, gloc)
  ADGA_OP_SWITCH(F)
#undef F
  return true;
}

bool launch_op_ndga(const Geom& g, const z64* hk, const z64* siw, const z64* dc, const int* keom, const z64* ssum,
                    z64* ndk, z64* nd, hipStream_t s) {
#define F(NN)                                                                                                   \
  hipLaunchKernelGGL((k_op_wsum<NN, true>), dim3(g.nkp_eom), dim3(OP_NT), 0, s, g, g.nkp_eom, hk, siw, dc, keom, \
                     ssum, ndk)
  ADGA_OP_SWITCH(F)
#undef F
  hipLaunchKernelGGL(k_op_ndga_mean, dim3(g.n), dim3(OP_NT), 0, s, g.n, g.nkp_eom, ndk, nd);
  return true;
}

}  // namespace adga
