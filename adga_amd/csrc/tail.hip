// The bubble over frequencies beyond the BSE box, in FP64.
//
// Large range (the reference's debug option "Bubble", config_module.f90:467-477): bubble_loc and
// bubble_nl sum chi0 over nu in [lo, hi] = [-iwmax + iwbmax_small, iwmax - iwbmax_small - 1], every
// frequency for which G(nu - omega) exists (main.F90:513-520, 644-663); the BSE still uses the box.
//   * the local part: a wave per diagonal element, -beta g_i(nu) g_j(nu - omega) / beta^2 summed over nu;
//   * the q-dependent part outside the box: for all q of an omega at once as one k-grid convolution summed
//     over the outer frequencies in real space, P(R) = sum_nu Ghat_il(nu, -R) Ghat_k'j(nu - omega, R), then
//     one inverse transform (the same convolution theorem as the per-nu bubble of kconv.hip).
//
// Matsubara tail (no counterpart in the reference): the local bubble beyond [lo, hi], nu_n with n >= K or
// n < -K (K = hi + 1), from the asymptotic expansion g(nu) = 1/x + c1/x^2 + c2/x^3 + O(x^-4), x = i nu:
//   -1/beta sum g_i(nu) g_j(nu - omega) = -1/beta [S11 + c1_i S21 + c1_j S12 + c2_i S31 + c1_i c1_j S22
//                                                   + c2_j S13] + O(K^-5)
// with S_pr = sum x^-p y^-r, y = i (nu - omega_m).  S11 in closed form: -2 (beta/2pi)^2 psi_1(K + 1/2) at
// m = 0 (psi_1: trigamma, asymptotic series of Abramowitz & Stegun 6.4.12 after upward recurrence), and
// -1/|omega_m| sum_{j=K-|m|}^{K+|m|-1} 1/nu_j otherwise (1/(x y) = (1/y - 1/x)/(x - y) telescopes).  The
// others by direct summation: S_pr = (-i)^(p+r) R_pr with the real pair sums
// R_pr = sum_{n=K}^{K+L-1} [nu_n^-p (nu_n - omega)^-r + (-nu_n)^-p (-nu_n - omega)^-r], whose terms are
// O(nu^-4), so the remainder beyond L = 2^18 is below 1e-15 of the sum.  The moments come from the two
// highest positive frequencies of giw: u = nu^2 Re g = -c1 + c3 / nu^2 and v = nu^3 (Im g + 1/nu) =
// c2 - c4 / nu^2 are linear in 1/nu^2 up to O(nu^-4).
#include "adga_kernels.h"

namespace adga {

namespace {

constexpr int TAIL_THREADS = 256;
constexpr long TAIL_L = 1L << 18;

__device__ double trigamma(double z) {
  double acc = 0.0;
  while (z < 20.0) {
    acc += 1.0 / (z * z);
    z += 1.0;
  }
  const double r = 1.0 / z, r2 = r * r;
  // 1/z + 1/(2 z^2) + sum_k B_2k / z^(2k+1), k = 1..7
  const double series =
      r * (1.0 + r * (0.5 + r * (1.0 / 6.0 + r2 * (-1.0 / 30.0 + r2 * (1.0 / 42.0 + r2 * (-1.0 / 30.0 +
      r2 * (5.0 / 66.0 + r2 * (-691.0 / 2730.0 + r2 * (7.0 / 6.0)))))))));
  return acc + series;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < TAIL_THREADS / 64; ++w) s += red[w];
  return s;
}

}  // namespace

// bubloc(i1, i1) = -1/beta sum_{nu in [lo, hi]} g_i(nu) g_j(nu - omega), i1 = i*n + j (one wave each);
// the off-diagonal entries are left as they are (zero).
__global__ void __launch_bounds__(64) k_bubloc_range(Geom g, int iwb, int lo, int hi, const z64* __restrict__ giw,
                                                     z64* __restrict__ bubloc) {
  const int i1 = blockIdx.x, lane = threadIdx.x;
  const int i = i1 / g.n, j = i1 % g.n;
  z64 acc = mk(0.0, 0.0);
  for (int w = lo + lane; w <= hi; w += 64) {
    const z64 gi = giw[(w + g.iwmax) + (long)i * 2 * g.iwmax];
    const z64 gj = giw[(w - iwb + g.iwmax) + (long)j * 2 * g.iwmax];
    acc = cfma(gi, gj, acc);
  }
  for (int o = 32; o > 0; o >>= 1) {
    acc.re += __shfl_down(acc.re, o, 64);
    acc.im += __shfl_down(acc.im, o, 64);
  }
  if (lane == 0) bubloc[i1 + (long)i1 * g.n2] = cscale(acc, -1.0 / g.beta);
}

void launch_bubloc_range(const Geom& g, int iwb, int lo, int hi, const z64* giw, z64* bubloc, hipStream_t s) {
  hipLaunchKernelGGL(k_bubloc_range, dim3(g.n2), dim3(64), 0, s, g, iwb, lo, hi, giw, bubloc);
}

// out = bubloc + the tail beyond nu_n, n in [-K, K-1], on the diagonal; one workgroup.
__global__ void __launch_bounds__(TAIL_THREADS) k_tail(Geom g, int iwb, int K, const z64* __restrict__ giw,
                                                       const z64* __restrict__ bubloc, z64* __restrict__ out) {
  __shared__ double red[TAIL_THREADS / 64];
  __shared__ double R[5];
  const int t = threadIdx.x;
  const double pb = M_PI / g.beta;
  const double om = 2.0 * pb * iwb;
  // pair sums R21, R12, R31, R22, R13
  double r21 = 0.0, r12 = 0.0, r31 = 0.0, r22 = 0.0, r13 = 0.0;
  for (long n = K + t; n < K + TAIL_L; n += TAIL_THREADS) {
    const double nu = pb * (2.0 * n + 1.0);
    const double a = 1.0 / nu, b = 1.0 / (nu - om), bp = -1.0 / (nu + om);  // 1/nu, 1/(nu-om), 1/(-nu-om)
    const double a2 = a * a;
    // (-nu)^-p = (-1)^p nu^-p
    r21 += a2 * b + a2 * bp;
    r12 += a * b * b - a * bp * bp;
    r31 += a2 * a * b - a2 * a * bp;
    r22 += a2 * b * b + a2 * bp * bp;
    r13 += a * b * b * b - a * bp * bp * bp;
  }
  r21 = block_sum(r21, red);
  r12 = block_sum(r12, red);
  r31 = block_sum(r31, red);
  r22 = block_sum(r22, red);
  r13 = block_sum(r13, red);
  if (t == 0) {
    R[0] = r21; R[1] = r12; R[2] = r31; R[3] = r22; R[4] = r13;
  }
  __syncthreads();
  // S11 in closed form
  double s11;
  if (iwb == 0) {
    const double c = g.beta / (2.0 * M_PI);
    s11 = -2.0 * c * c * trigamma(K + 0.5);
  } else {
    const int M = iwb < 0 ? -iwb : iwb;
    double h = 0.0;
    for (int jj = K - M; jj < K + M; ++jj) h += 1.0 / (pb * (2.0 * jj + 1.0));
    s11 = -h / (2.0 * pb * M);
  }
  // S_pr = (-i)^(p+r) R_pr: (-i)^3 = i, (-i)^4 = 1
  const z64 S21 = mk(0.0, R[0]), S12 = mk(0.0, R[1]);
  const double S31 = R[2], S22 = R[3], S13 = R[4];
  for (int e = t; e < g.n4; e += TAIL_THREADS) {
    const int a1 = e % g.n2, a2 = e / g.n2;
    z64 v = bubloc[e];
    if (a1 == a2) {
      const int i = a1 / g.n, j = a1 % g.n;
      double c1[2], c2[2];
      for (int side = 0; side < 2; ++side) {
        const int orb = side == 0 ? i : j;
        const double n0 = pb * (2.0 * (g.iwmax - 1) + 1.0), n1 = pb * (2.0 * (g.iwmax - 2) + 1.0);
        const z64 g0 = giw[(2 * g.iwmax - 1) + (long)orb * 2 * g.iwmax];
        const z64 g1 = giw[(2 * g.iwmax - 2) + (long)orb * 2 * g.iwmax];
        const double h0 = 1.0 / (n0 * n0), h1 = 1.0 / (n1 * n1);
        const double u0 = n0 * n0 * g0.re, u1 = n1 * n1 * g1.re;
        const double v0 = n0 * n0 * n0 * (g0.im + 1.0 / n0), v1 = n1 * n1 * n1 * (g1.im + 1.0 / n1);
        c1[side] = -(u0 * h1 - u1 * h0) / (h1 - h0);
        c2[side] = (v0 * h1 - v1 * h0) / (h1 - h0);
      }
      z64 s = mk(s11 + c2[0] * S31 + c1[0] * c1[1] * S22 + c2[1] * S13, 0.0);
      s = cadd(s, cadd(cscale(S21, c1[0]), cscale(S12, c1[1])));
      v = cadd(v, cscale(s, -1.0 / g.beta));
    }
    out[e] = v;
  }
}

void launch_tail(const Geom& g, int iwb, int K, const z64* giw, const z64* bubloc, z64* out, hipStream_t s) {
  hipLaunchKernelGGL(k_tail, dim3(1), dim3(TAIL_THREADS), 0, s, g, iwb, K, giw, bubloc, out);
}

// P[c][R] = sum over nu in [lo, hi] outside the box [-g.iwf, g.iwf) of Ghat(nu)_il(-R) Ghat(nu - omega)_k'j(R),
// c = ((i*n + l)*n + k')*n + j; ghat covers nu in [g.glo, g.glo + g.nw).
__global__ void __launch_bounds__(256) k_bub_outer_prod(Geom g, int iwb, int lo, int hi,
                                                        const z64* __restrict__ ghat, z64* __restrict__ P) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)g.n4 * g.nkp) return;
  const int R = (int)(idx % g.nkp);
  const int c = (int)(idx / g.nkp);
  const int il = c / g.n2, kj = c % g.n2;
  const int rz = R % g.nkz, ry = (R / g.nkz) % g.nky, rx = R / (g.nky * g.nkz);
  const int mR = ((g.nkz - rz) % g.nkz) + ((g.nky - ry) % g.nky) * g.nkz + ((g.nkx - rx) % g.nkx) * g.nky * g.nkz;
  const long fs = (long)g.n2 * g.nkp;  // one frequency's fields
  const z64* a = ghat + (long)il * g.nkp + mR;
  const z64* b = ghat + (long)kj * g.nkp + R;
  z64 acc = mk(0.0, 0.0);
  for (int w = lo; w <= hi; ++w) {
    if (w == -g.iwf) {
      w = g.iwf - 1;  // skip the box
      continue;
    }
    acc = cfma(a[(long)(w - g.glo) * fs], b[(long)(w - iwb - g.glo) * fs], acc);
  }
  P[idx] = acc;
}

void launch_bubble_outer(const Geom& g, int iwb, int lo, int hi, const z64* tw, const z64* ghat, z64* P, z64* out,
                         hipStream_t s) {
  const long tot = (long)g.n4 * g.nkp;
  hipLaunchKernelGGL(k_bub_outer_prod, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, iwb, lo, hi, ghat,
                     P);
  dft3(g, g.n4, -1, tw, P, out, s);
}

// bub_out[b](i1 = i*n + j, i2 = l*n + k') += -1/(beta nkp^2) all[c][q_b]: the outer frequencies' part of
// sum_nu chi0^q / beta^2 (chi0 = -beta/nkp sum_k, and 1/nkp from the convolution theorem)
__global__ void k_bub_outer_add(Geom g, int npts, const int* __restrict__ qidx, const z64* __restrict__ all,
                                z64* __restrict__ bub_out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npts * g.n4) return;
  const int c = (int)(idx % g.n4);
  const int b = (int)(idx / g.n4);
  const int i = c / (g.n * g.n2), l = (c / g.n2) % g.n, kp = (c / g.n) % g.n, j = c % g.n;
  const double f = -1.0 / (g.beta * (double)g.nkp * g.nkp);
  z64* o = bub_out + (long)b * g.n4 + (i * g.n + j) + (long)(l * g.n + kp) * g.n2;
  *o = cadd(*o, cscale(all[(long)c * g.nkp + qidx[b]], f));
}

void launch_bubble_outer_add(const Geom& g, int npts, const int* qidx, const z64* all, z64* bub_out, hipStream_t s) {
  const long tot = (long)npts * g.n4;
  hipLaunchKernelGGL(k_bub_outer_add, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, qidx, all,
                     bub_out);
}

}  // namespace adga
