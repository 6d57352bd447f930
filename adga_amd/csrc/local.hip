// A1 (G(k,nu) table) and A4/A5 (local bubble, chi0^w F, gamma^w) on gfx950.
#include "adga_kernels.h"
#include "small_inv.h"

namespace adga {

// ---------------------------------------------------------------- A1: G table
// Gt[(w*nkp + k)*n2 + a*n + b] = G(k, nu_{glo+w})_{ab}
//   G^-1 = (i nu + mu) - H(k) - dc - Sigma(nu) on the diagonal, -H(k) off it
//   (one_particle_quant_module.f90:240-271, get_gkiw 81-104).
template <int NN>
__global__ void __launch_bounds__(256) k_gtable(Geom g, const z64* __restrict__ hk, const z64* __restrict__ siw,
                                                const z64* __restrict__ dc, z64* __restrict__ gt) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)g.nkp * g.nw) return;
  int k = (int)(idx % g.nkp);
  int w = (int)(idx / g.nkp);
  int iw = g.glo + w;
  double nu = M_PI * (2 * iw + 1) / g.beta;
  z64 a[NN][NN];
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      z64 h = hk[(long)k * NN * NN + i + j * NN];  // hk(i, j, k), Fortran order
      a[i][j] = mk(-h.re, -h.im);
    }
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    z64 h = hk[(long)k * NN * NN + i + i * NN];
    z64 s = siw[(iw + g.iwmax) + (long)i * 2 * g.iwmax];
    z64 d = dc[2 * i];  // dc(1, i)
    a[i][i] = mk(g.mu - h.re - d.re - s.re, nu - h.im - d.im - s.im);
  }
  inv_small<NN>(a);
  z64* out = gt + idx * (NN * NN);
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int j = 0; j < NN; ++j) out[i * NN + j] = a[i][j];
}

void launch_gtable(const Geom& g, const z64* hk, const z64* siw, const z64* dc, z64* gt, hipStream_t s) {
  long tot = (long)g.nkp * g.nw;
  dim3 grid((unsigned)((tot + 255) / 256));
  switch (g.n) {
    case 1: hipLaunchKernelGGL(k_gtable<1>, grid, dim3(256), 0, s, g, hk, siw, dc, gt); break;
    case 2: hipLaunchKernelGGL(k_gtable<2>, grid, dim3(256), 0, s, g, hk, siw, dc, gt); break;
    case 3: hipLaunchKernelGGL(k_gtable<3>, grid, dim3(256), 0, s, g, hk, siw, dc, gt); break;
    case 4: hipLaunchKernelGGL(k_gtable<4>, grid, dim3(256), 0, s, g, hk, siw, dc, gt); break;
    case 5: hipLaunchKernelGGL(k_gtable<5>, grid, dim3(256), 0, s, g, hk, siw, dc, gt); break;
    case 6: hipLaunchKernelGGL(k_gtable<6>, grid, dim3(256), 0, s, g, hk, siw, dc, gt); break;
    case 7: hipLaunchKernelGGL(k_gtable<7>, grid, dim3(256), 0, s, g, hk, siw, dc, gt); break;
    case 8: hipLaunchKernelGGL(k_gtable<8>, grid, dim3(256), 0, s, g, hk, siw, dc, gt); break;
    default: break;
  }
}

// ---------------------------------------------------------------- A4: local bubble
// winv(s, d) = -1/(beta g_i(nu) g_j(nu-omega)), chi0w(s, d) = -beta g_i g_j, s = i*n + j
// (get_chi0_loc 192-207, get_chi0_loc_inv 210-225).
__global__ void k_local_bubble(Geom g, int iwb, const z64* __restrict__ giw, z64* __restrict__ winv,
                               z64* __restrict__ chi0w) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.n2 * g.nf) return;
  int s = idx % g.n2, d = idx / g.n2;
  int i = s / g.n, j = s % g.n;
  int iwf = d - g.iwf;
  z64 gi = giw[(iwf + g.iwmax) + (long)i * 2 * g.iwmax];
  z64 gj = giw[(iwf - iwb + g.iwmax) + (long)j * 2 * g.iwmax];
  z64 bg = cmul(mk(g.beta * gi.re, g.beta * gi.im), gj);
  chi0w[idx] = mk(-bg.re, -bg.im);
  z64 r = crecip(bg);
  winv[idx] = mk(-r.re, -r.im);
}

// ---------------------------------------------------------------- A5: chi0^w F
// C(i1, i2) = 0.5 (X(i1,i2) + X(i2,i1)) * winv(i2) - delta(i1,i2), Fortran column-major
// (main.F90:529-553).  32x32 tiles through LDS so both X reads are coalesced.
__global__ void __launch_bounds__(256) k_chi0wF(int N, const z64* __restrict__ x, const z64* __restrict__ winv,
                                                z64* __restrict__ c) {
  __shared__ z64 tile[32][33];
  int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx: i1 block, by: i2 block
  int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 8 rows of threads
  // tile[a][b] = X(by + a, bx + b), loaded with the first (contiguous) index along the threads
  for (int r = ty; r < 32; r += 8) {
    int row = by + tx, col = bx + r;
    if (row < N && col < N) tile[tx][r] = x[(long)row + (long)col * N];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    int i1 = bx + tx, i2 = by + r;
    if (i1 < N && i2 < N) {
      z64 a = x[(long)i1 + (long)i2 * N];
      z64 b = tile[r][tx];
      z64 sym = mk(0.5 * (a.re + b.re), 0.5 * (a.im + b.im));
      z64 v = cmul(sym, winv[i2]);
      if (i1 == i2) v.re -= 1.0;
      c[(long)i1 + (long)i2 * N] = v;
    }
  }
}

// gamma^w(s, j) = sum_d C(d*n2 + s, j); (1+gamma^w)(s, j) = gamma^w(s, j) + [j % n2 == s]
// (main.F90:557-571, 616-625).  Output (n2, N) Fortran.
__global__ void k_gammaw(Geom g, const z64* __restrict__ c, z64* __restrict__ gam, z64* __restrict__ opg) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)g.n2 * g.N) return;
  int s = (int)(idx % g.n2);
  long j = idx / g.n2;
  z64 acc = mk(0.0, 0.0);
  for (int d = 0; d < g.nf; ++d) acc = cadd(acc, c[(long)(d * g.n2 + s) + j * g.N]);
  gam[idx] = acc;
  if ((int)(j % g.n2) == s) acc.re += 1.0;
  opg[idx] = acc;
}

// Local chi (calc_chi_qw(gamma^w, chi0w), main.F90:573-577; susc_module.f90:34-56) and the local
// bubble sum_nu chi0w / beta^2 (main.F90:512-519).  One thread per output (s1, s2).
__global__ void k_local_chi(Geom g, const z64* __restrict__ gam_d, const z64* __restrict__ gam_m,
                            const z64* __restrict__ chi0w, z64* __restrict__ out_d, z64* __restrict__ out_m,
                            z64* __restrict__ bubloc) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.n2 * g.n2) return;
  int s1 = idx % g.n2, s2 = idx / g.n2;
  z64 ad = mk(0.0, 0.0), am = mk(0.0, 0.0), bl = mk(0.0, 0.0);
  double ib2 = 1.0 / (g.beta * g.beta);
  for (int d = 0; d < g.nf; ++d) {
    z64 w = chi0w[s2 + d * g.n2];  // diagonal element chi0w(s2, s2, d)
    long col = (long)(d * g.n2 + s2) * g.n2;
    ad = cfma(gam_d[s1 + col], w, ad);
    am = cfma(gam_m[s1 + col], w, am);
    if (s1 == s2) bl = cadd(bl, w);
  }
  out_d[idx] = cscale(ad, ib2);
  out_m[idx] = cscale(am, ib2);
  bubloc[idx] = cscale(bl, ib2);
}

void launch_local(const Geom& g, int iwb, const z64* giw, const z64* xd, const z64* xm, z64* winv, z64* chi0w,
                  z64* cd, z64* cm, z64* gam_d, z64* gam_m, z64* opg_d, z64* opg_m, z64* chiloc_d,
                  z64* chiloc_m, z64* bubloc, hipStream_t s) {
  int nb = (g.n2 * g.nf + 255) / 256;
  hipLaunchKernelGGL(k_local_bubble, dim3(nb), dim3(256), 0, s, g, iwb, giw, winv, chi0w);
  dim3 tg((g.N + 31) / 32, (g.N + 31) / 32);
  hipLaunchKernelGGL(k_chi0wF, tg, dim3(256), 0, s, g.N, xd, winv, cd);
  hipLaunchKernelGGL(k_chi0wF, tg, dim3(256), 0, s, g.N, xm, winv, cm);
  long tot = (long)g.n2 * g.N;
  dim3 gg((unsigned)((tot + 255) / 256));
  hipLaunchKernelGGL(k_gammaw, gg, dim3(256), 0, s, g, cd, gam_d, opg_d);
  hipLaunchKernelGGL(k_gammaw, gg, dim3(256), 0, s, g, cm, gam_m, opg_m);
  hipLaunchKernelGGL(k_local_chi, dim3((g.n4 + 255) / 256), dim3(256), 0, s, g, gam_d, gam_m, chi0w, chiloc_d,
                     chiloc_m, bubloc);
}

}  // namespace adga
