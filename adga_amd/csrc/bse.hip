// A3 (bubble), A6 (BSE build) and the FP64 right-hand side / residuals of the A7/A8 solves
// (the LU itself is in lu.hip) on gfx950.
//
// Formulation (main.F90:673-757): for each channel r the reference forms
// A = 1 - blockdiag(S_nu) C with S_nu = chi0^q(nu) diag(chi0^w(nu))^-1 - 1 and
// C = chi0^w F_r, inverts it, and forms eta = (1+gamma^w)(A^-1 - 1).  Here the
// same eta is obtained without an explicit inverse and without the cancellation
// in A^-1 - 1: eta A = (1+gamma^w)(1 - A) = (1+gamma^w) blockdiag(S) C is solved
// as A^T eta^T = B by LU with partial pivoting of M = A^T, with the n^2
// right-hand sides appended as extra columns of M so the forward elimination
// happens inside the factorisation.  B and every residual are formed in FP64
// from the unrounded S and C, so the residual is on the scale of eta itself.
// M is stored row-major per matrix (row r, column c at r*ldm + c).
#include "adga_kernels.h"
#include "mfma_gemm.h"

namespace adga {

// ---------------------------------------------------------------- A3: bubble
// chi0q[b][d](i1 = i*n + j, i2 = l*n + k') = -beta/nkp sum_K G(K, nu_d)_il G(K - q_b, nu_d - omega)_k'j
// (accumulate_chi0, one_particle_quant_module.f90:274-288; k-sum main.F90:647-657).
// One block per (nu, point); thread = ((i,l), k-slice); n^2 accumulators per thread.
// gt2 is the table of the second factor G(K - q, nu - omega) (gt itself, or a derivative table when the
// tangent of the bubble is formed by the product rule); accumulate adds into chi0q.
template <int NN>
__global__ void __launch_bounds__(256) k_bubble(Geom g, int iwb, const int* __restrict__ qidx,
                                                const z64* __restrict__ gt, const z64* __restrict__ gt2,
                                                z64* __restrict__ chi0q, int accumulate) {
  constexpr int N2 = NN * NN;
  constexpr int NKS = 256 / N2;
  __shared__ z64 red[256];
  const int d = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const int og = t % N2, ks = t / N2;
  const int i = og / NN, l = og % NN;
  const int q = qidx[b];
  const int iwf = d - g.iwf;
  const z64* g1 = gt + (long)(iwf - g.glo) * g.nkp * N2;
  const z64* g2 = gt2 + (long)(iwf - iwb - g.glo) * g.nkp * N2;
  z64 acc[N2];
#pragma unroll
  for (int e = 0; e < N2; ++e) acc[e] = mk(0.0, 0.0);
  if (ks < NKS) {
    for (int K = ks; K < g.nkp; K += NKS) {
      z64 a = g1[(long)K * N2 + i * NN + l];
      int kq = k_minus_q(K, q, g.nkx, g.nky, g.nkz);
      const z64* row = g2 + (long)kq * N2;
#pragma unroll
      for (int e = 0; e < N2; ++e) acc[e] = cfma(a, row[e], acc[e]);  // e = k'*n + j
    }
  }
  const double f = -g.beta / g.nkp;
  z64* out = chi0q + ((long)b * g.nf + d) * g.n4;
#pragma unroll
  for (int e = 0; e < N2; ++e) {
    red[t] = acc[e];
    __syncthreads();
    if (t < N2) {
      z64 s = mk(0.0, 0.0);
      for (int kk = 0; kk < NKS; ++kk) s = cadd(s, red[kk * N2 + t]);
      // thread t = og = (i, l); e = (k', j): chi0(i1 = i*n + j, i2 = l*n + k')
      int ii = t / NN, ll = t % NN, kp = e / NN, jj = e % NN;
      z64* o = out + (ii * NN + jj) + (ll * NN + kp) * N2;
      *o = accumulate ? cadd(*o, cscale(s, f)) : cscale(s, f);
    }
    __syncthreads();
  }
}

void launch_bubble(const Geom& g, int iwb, int npts, const int* qidx, const z64* gt, z64* chi0q, hipStream_t s,
                   const z64* gt2, bool accumulate) {
  dim3 grid(g.nf, npts);
  const z64* g2 = gt2 ? gt2 : gt;
  const int acc = accumulate ? 1 : 0;
  switch (g.n) {
    case 1: hipLaunchKernelGGL(k_bubble<1>, grid, dim3(256), 0, s, g, iwb, qidx, gt, g2, chi0q, acc); break;
    case 2: hipLaunchKernelGGL(k_bubble<2>, grid, dim3(256), 0, s, g, iwb, qidx, gt, g2, chi0q, acc); break;
    case 3: hipLaunchKernelGGL(k_bubble<3>, grid, dim3(256), 0, s, g, iwb, qidx, gt, g2, chi0q, acc); break;
    case 4: hipLaunchKernelGGL(k_bubble<4>, grid, dim3(256), 0, s, g, iwb, qidx, gt, g2, chi0q, acc); break;
    case 5: hipLaunchKernelGGL(k_bubble<5>, grid, dim3(256), 0, s, g, iwb, qidx, gt, g2, chi0q, acc); break;
    default: break;
  }
}

// ---------------------------------------------------------------- S_nu
// S[b][d][a][c] = chi0q(a, c, d) winv(c, d) - delta_ac  (main.F90:685-688)
__global__ void k_smat(Geom g, int npts, const z64* __restrict__ chi0q, const z64* __restrict__ winv,
                       z64* __restrict__ S) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)npts * g.nf * g.n4;
  if (idx >= tot) return;
  int e = (int)(idx % g.n4);
  long bd = idx / g.n4;
  int d = (int)(bd % g.nf);
  int a = e / g.n2, c = e % g.n2;  // S row a, column c (row-major)
  z64 v = cmul(chi0q[bd * g.n4 + a + (long)c * g.n2], winv[c + d * g.n2]);
  if (a == c) v.re -= 1.0;
  S[idx] = v;
}

void launch_smat(const Geom& g, int npts, const z64* chi0q, const z64* winv, z64* S, hipStream_t s) {
  long tot = (long)npts * g.nf * g.n4;
  hipLaunchKernelGGL(k_smat, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, chi0q, winv, S);
}

// ---------------------------------------------------------------- A6: BSE build
// M_{b,ch}(r, d*n2 + s) = delta - R_d(s, r),  R_d(s, r) = sum_i' S_b[d][s][i'] C_ch(d*n2 + i', r)
// (main.F90:693-700, transposed: M = A^T).  A workgroup owns DB frequency blocks d of one channel
// for PB points and all rows: their S_d sit in LDS, and each thread reads the n2 contiguous values
// C(d*n2 + ., r) (a row slice of C^T) once and turns them into n2 outputs per point.  Padding
// columns [N, ldm) are not written (never read).  The complex64 build (the refined path's LU
// input, whose residuals are formed in FP64 from S and C themselves) computes in FP32: its entries
// then differ from the rounded FP64 ones by a few FP32 ulps.  The complex128 build computes in FP64.
// Frequency blocks per workgroup (DB) and points per workgroup (PB, sharing each C row read).  For
// n <= 3 a workgroup's row segment is DB * n2 elements: complex64 takes 16 blocks, so that at n = 3 each
// segment is 1152 bytes and starts on a 128-byte line (8 blocks gave 576-byte segments, every other one
// starting mid-line; 21.2 -> 20.3 ms per build of 8000 matrices at N = 1080), with 2 points per
// workgroup to keep four workgroups per CU; complex128 segments of 8 blocks are already whole lines.
template <int NN, int ES>
struct BuildShape {
  static constexpr int N2 = NN * NN, N4 = N2 * N2;
  static constexpr int DB = NN <= 3 ? (ES == 8 ? 16 : 8) : 2;
  static constexpr int PB0 = 40960 / (DB * N4 * ES);
  static constexpr int PB1 = PB0 < 1 ? 1 : PB0 > 4 ? 4 : PB0;
  static constexpr int PB = NN <= 3 ? ((ES == 8 && PB1 > 2) ? 2 : PB1) : (PB1 < (ES == 8 ? 2 : 1) ? PB1 : (ES == 8 ? 2 : 1));
};

template <typename T, int NN>
__global__ void __launch_bounds__(256) k_build(Geom g, int npts, int ldm, const z64* __restrict__ S,
                                               const z64* __restrict__ c_d, const z64* __restrict__ c_m,
                                               cplx<T>* __restrict__ M) {
  using CA = cplx<T>;  // arithmetic type
  using Sh = BuildShape<NN, (int)sizeof(CA)>;
  constexpr int N2 = Sh::N2, N4 = Sh::N4, DB = Sh::DB, PB = Sh::PB;
  __shared__ CA sS[PB][DB * N4];
  // grid (point groups, channels, frequency blocks): the point group runs fastest, so the workgroups
  // resident at one time read the same column block of C, which stays in their XCDs' L2
  const int ch = blockIdx.y, b0 = blockIdx.x * PB, t = threadIdx.x;
  const int N = g.N;
  const int d0 = blockIdx.z * DB;
  const int nd = min(DB, g.nf - d0);
  const int np = min(PB, npts - b0);
  for (int pb = 0; pb < np; ++pb) {
    const z64* Sb = S + ((long)(b0 + pb) * g.nf + d0) * N4;
    for (int e = t; e < nd * N4; e += 256) sS[pb][e] = ccast<T>(Sb[e]);
  }
  __syncthreads();
  const z64* C = ch == 0 ? c_d : c_m;
  // outputs go through an LDS tile so that each row segment (nd*n2 complex) is written by consecutive
  // lanes.  QT threads share a (row, frequency block) and split its n2 outputs: one for n <= 3, four
  // for n = 4, 5 (16 or 25 outputs per row and block), whose tile would otherwise span 128 rows
  // (52 KB of complex64 at n = 5) and leave few workgroups per CU
  constexpr int QT = NN <= 3 ? 1 : 4, QS = (N2 + QT - 1) / QT;
  constexpr int RS = 256 / (DB * QT), W = DB * N2;
  const int q = t % QT, dl = (t / QT) % DB, rl = t / (QT * DB);
  const int d = d0 + dl, sq0 = q * QS, sq1 = min(N2, sq0 + QS);
  __shared__ cplx<T> tile[RS][W + 1];
  for (int r0 = 0; r0 < N; r0 += RS) {
    const int r = r0 + rl;
    const bool act = dl < nd && r < N;
    CA cc[N2];
    const z64* crow = C + (long)(act ? r : 0) * N + (act ? d : 0) * N2;
#pragma unroll
    for (int i = 0; i < N2; ++i) cc[i] = ccast<T>(crow[i]);
    for (int pb = 0; pb < np; ++pb) {
      if (act) {
        const CA* sd = sS[pb] + dl * N4;
        // not unrolled: S_d stays in LDS (hoisting its n2^2 values into registers costs occupancy)
#pragma unroll 1
        for (int sq = sq0; sq < sq1; ++sq) {
          CA acc = mk<T>(0, 0);
#pragma unroll
          for (int i = 0; i < N2; ++i) acc = cfma(sd[sq * N2 + i], cc[i], acc);
          tile[rl][dl * N2 + sq] = mk<T>((d * N2 + sq == r ? (T)1 : (T)0) - acc.re, -acc.im);
        }
      }
      __syncthreads();
      // 16-byte stores: ldm is a multiple of 16 and d0 * n2 is even, so every row segment starts
      // 16-byte aligned; non-temporal, as M is read again only by the factorisation after the whole
      // build (20.6 -> 19.2 ms per complex64 build of 8000 matrices at N = 1080, 34.8 -> 33.1 complex128)
      cplx<T>* mb = M + ((long)(2 * (b0 + pb) + ch) * N + r0) * ldm + d0 * N2;
      constexpr int VW = 16 / (int)sizeof(cplx<T>), WV = W / VW;
      static_assert(W % VW == 0, "row segment of whole 16-byte vectors");
      struct alignas(16) Vec { cplx<T> v[VW]; };
      const int lim = nd * N2;
      for (int e = t; e < RS * WV; e += 256) {
        const int row = e / WV, col = (e % WV) * VW;
        if (r0 + row >= N) continue;
        cplx<T>* dst = mb + (long)row * ldm + col;
        if (col + VW <= lim) {
          Vec x;
#pragma unroll
          for (int j = 0; j < VW; ++j) x.v[j] = tile[row][col + j];
          if constexpr (sizeof(T) == 4) {
            typedef float f4 __attribute__((ext_vector_type(4)));
            const f4 vv = {x.v[0].re, x.v[0].im, x.v[1].re, x.v[1].im};
            __builtin_nontemporal_store(vv, reinterpret_cast<f4*>(dst));
          } else {
            typedef double d2 __attribute__((ext_vector_type(2)));
            const d2 vv = {x.v[0].re, x.v[0].im};
            __builtin_nontemporal_store(vv, reinterpret_cast<d2*>(dst));
          }
        } else {
          for (int j = 0; j < VW; ++j)
            if (col + j < lim) dst[j] = tile[row][col + j];
        }
      }
      __syncthreads();
    }
  }
}

template <typename T>
void launch_build(const Geom& g, int npts, int ldm, const z64* S, const z64* cd, const z64* cm, cplx<T>* M,
                  hipStream_t s) {
#define ADGA_BUILD_CASE(NN)                                                                            \
  case NN: {                                                                                           \
    using Sh = BuildShape<NN, (int)sizeof(cplx<T>)>;                                                   \
    constexpr int DB = Sh::DB, PB = Sh::PB;                                                           \
    dim3 grid((npts + PB - 1) / PB, 2, (g.nf + DB - 1) / DB);                                         \
    hipLaunchKernelGGL((k_build<T, NN>), grid, dim3(256), 0, s, g, npts, ldm, S, cd, cm, M);          \
    break;                                                                                             \
  }
  switch (g.n) {
    ADGA_BUILD_CASE(1)
    ADGA_BUILD_CASE(2)
    ADGA_BUILD_CASE(3)
    ADGA_BUILD_CASE(4)
    ADGA_BUILD_CASE(5)
    default: break;
  }
#undef ADGA_BUILD_CASE
}
template void launch_build<float>(const Geom&, int, int, const z64*, const z64*, const z64*, c32*, hipStream_t);
template void launch_build<double>(const Geom&, int, int, const z64*, const z64*, const z64*, z64*, hipStream_t);

// ---------------------------------------------------------------- A7/A8 refinement residuals
// The residual products take their right operand row-major, one N x (npts n2) plane per channel:
// Z_ch[x][b*n2 + s] = sum_c' S_b[d][c'][i'] V[2b + ch][d*n2 + c'][s]  (x = d*n2 + i'), i.e.
// Z = blockdiag(S)^T V with the points side by side, V = Y (a solution) or, when Y is null,
// V = (1 + gamma^w_ch)^T.  Threads run over one matrix m = 2b + ch at a time (x, then s fastest),
// so the reads of S_b and V_m stay local; the writes land in runs of n2 values.
__global__ void k_resid_z(Geom g, int npts, const z64* __restrict__ S, const z64* __restrict__ Y,
                          const z64* __restrict__ opg_d, const z64* __restrict__ opg_m, z64* __restrict__ Z) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ncol = (long)npts * g.n2;
  if (idx >= 2 * g.N * ncol) return;
  const int s = (int)(idx % g.n2);
  const long mx = idx / g.n2;
  const int x = (int)(mx % g.N), m = (int)(mx / g.N);
  const int b = m / 2, ch = m % 2;
  const int d = x / g.n2, ip = x % g.n2;
  const z64* Sd = S + ((long)b * g.nf + d) * g.n4;
  z64 acc = mk(0.0, 0.0);
  if (Y) {
    const z64* Ym = Y + (long)m * g.N * g.n2;
    for (int cp = 0; cp < g.n2; ++cp) acc = cfma(Sd[cp * g.n2 + ip], Ym[(long)(d * g.n2 + cp) * g.n2 + s], acc);
  } else {
    const z64* opg = ch == 0 ? opg_d : opg_m;
    for (int cp = 0; cp < g.n2; ++cp) acc = cfma(Sd[cp * g.n2 + ip], opg[s + (long)(d * g.n2 + cp) * g.n2], acc);
  }
  Z[((long)ch * g.N + x) * ncol + (long)b * g.n2 + s] = acc;
}

// Out[(2b + ch) N + r][s] = sum_x C_ch(x, r) Z_ch[x][b*n2 + s], plus (B - Y) at the same position
// when B is given (the residual): the FP64 products on the 3M tile (128 x 64, eight waves), C read
// row-major as A(r, x) = C[r N + x], the output in the point-major layout of Y.  Workgroups are
// dealt round-robin to the 8 XCDs; the decode puts all row tiles of one column tile on one XCD,
// adjacent in time, so that the column block of Z (N x 64) is shared through its L2.
constexpr int RG_TM = 128, RG_TN = 64, RG_WR = 4, RG_WC = 2, RG_KC = 16;
__global__ void __launch_bounds__(RG_WR * RG_WC * 64)
k_resid_gemm(Geom g, int npts, int ch0, const z64* __restrict__ cd, const z64* __restrict__ cm,
             const z64* __restrict__ Z, z64* Out, const z64* __restrict__ B,
             const z64* __restrict__ Y, int accumulate) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[G3Cfg<double, RG_TM, RG_TN, RG_WR, RG_WC, RG_KC>::LDS];
  const int ch = ch0 + (int)blockIdx.y;
  const z64* C = ch == 0 ? cd : cm;
  const int N = g.N, n2 = g.n2, ncol = npts * n2;
  const int ntr = (N + RG_TM - 1) / RG_TM, ntc = (ncol + RG_TN - 1) / RG_TN;
  const int L = (int)blockIdx.x, xcd = L & 7, j = L >> 3;
  const int tc = (j / ntr) * 8 + xcd, tr = j % ntr;
  if (tc >= ntc) return;
  const int rt = tr * RG_TM, ct = tc * RG_TN;
  auto at = [=](int r, int col) {
    const int b = col / n2, sq = col - b * n2;
    return ((long)(2 * b + ch) * N + r) * n2 + sq;
  };
  gemm3m_tile<double, RG_TM, RG_TN, RG_WR, RG_WC, RG_KC>(
      C, N, Z + (long)ch * N * ncol, ncol,
      [=](int r, int col) {
        const long i = at(r, col);
        return B ? csub(B[i], Y[i]) : Out[i];
      },
      [=](int r, int col, z64 v) { Out[at(r, col)] = v; }, N, ncol, N, B == nullptr && !accumulate, false, rt, ct,
      lds);
}

// Out = C_ch^T Z_ch (+ B - Y when B is given, + Out itself when accumulate)
void resid_gemm(const Geom& g, int npts, int ch0, int nch, const z64* cd, const z64* cm, const z64* Z, z64* Out,
                const z64* B, const z64* Y, hipStream_t s, bool accumulate) {
  const int ntr = (g.N + RG_TM - 1) / RG_TM, ntc = (npts * g.n2 + RG_TN - 1) / RG_TN;
  const int nwg = ntr * ((ntc + 7) / 8) * 8;
  hipLaunchKernelGGL(k_resid_gemm, dim3(nwg, nch), dim3(RG_WR * RG_WC * 64), 0, s, g, npts, ch0, cd, cm, Z, Out, B,
                     Y, accumulate ? 1 : 0);
}

// Per-matrix max |R| / max |B| (|.| = |re|+|im|), one workgroup per matrix; a zero right-hand
// side counts as converged when the residual is zero too.
__global__ void __launch_bounds__(256) k_resid_norm(Geom g, const z64* __restrict__ R, const z64* __restrict__ B,
                                                    double* __restrict__ rel) {
  __shared__ double sr[256], sb[256];
  const int m = blockIdx.x, t = threadIdx.x;
  const long tot = (long)g.N * g.n2;
  double mr = 0.0, mb = 0.0;
  for (long e = t; e < tot; e += 256) {
    mr = fmax(mr, cabs1(R[(long)m * tot + e]));
    mb = fmax(mb, cabs1(B[(long)m * tot + e]));
    double v = R[(long)m * tot + e].re + R[(long)m * tot + e].im;
    if (!isfinite(v)) mr = INFINITY;
  }
  sr[t] = mr;
  sb[t] = mb;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) { sr[t] = fmax(sr[t], sr[t + o]); sb[t] = fmax(sb[t], sb[t + o]); }
    __syncthreads();
  }
  if (t == 0) rel[m] = sb[0] > 0.0 ? sr[0] / sb[0] : (sr[0] > 0.0 ? INFINITY : 0.0);
}

// Zg[d*n2 + i'][b*n2 + s] = S_b[d][s][i'] (the channel-0 plane of the residual product's right
// operand), so that C^T Zg (k_resid_gemm) is
// gamma^q_d(s, r) = sum_d sum_i' S_b[d][s][i'] C_d(d*n2 + i', r) = sum_nu chi0^{nl,q} F^w_d
// (main.F90:693-700, the accumulation of rectanglework over nu).
__global__ void k_gamq_z(Geom g, int npts, const z64* __restrict__ S, z64* __restrict__ Z) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ncol = (long)npts * g.n2;
  if (idx >= g.N * ncol) return;
  const int col = (int)(idx % ncol), x = (int)(idx / ncol);
  const int b = col / g.n2, s = col - b * g.n2;
  const int d = x / g.n2, ip = x % g.n2;
  Z[idx] = S[((long)b * g.nf + d) * g.n4 + s * g.n2 + ip];
}

// k_resid_z and k_gamq_z for n <= 3 with one thread per (channel, d, point b, column s) forming the n2 rows
// x = d*n2 + i' of its column: each thread reads its values of V (or of S_b) once, and for each i' the
// wave writes a contiguous run of Z's row x (consecutive lanes hold consecutive columns b*n2 + s).  The
// same sums in the same order as the one-output-per-thread forms above.
template <int N2>
__global__ void __launch_bounds__(256) k_resid_zt(Geom g, int npts, const z64* __restrict__ S,
                                                  const z64* __restrict__ Y, const z64* __restrict__ opg_d,
                                                  const z64* __restrict__ opg_m, z64* __restrict__ Z) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ncol = (long)npts * N2;
  if (idx >= 2L * g.nf * ncol) return;
  const int col = (int)(idx % ncol);
  const int r = (int)(idx / ncol), d = r % g.nf, ch = r / g.nf;
  const int b = col / N2, s = col - b * N2;
  const z64* Sd = S + ((long)b * g.nf + d) * (N2 * N2);
  z64 v[N2];
  if (Y) {
    const z64* Ym = Y + (long)(2 * b + ch) * g.N * N2;
#pragma unroll
    for (int cp = 0; cp < N2; ++cp) v[cp] = Ym[(long)(d * N2 + cp) * N2 + s];
  } else {
    const z64* opg = ch == 0 ? opg_d : opg_m;
#pragma unroll
    for (int cp = 0; cp < N2; ++cp) v[cp] = opg[s + (long)(d * N2 + cp) * N2];
  }
  z64* Zc = Z + ((long)ch * g.N + (long)d * N2) * ncol + col;
#pragma unroll
  for (int ip = 0; ip < N2; ++ip) {
    z64 acc = mk(0.0, 0.0);
#pragma unroll
    for (int cp = 0; cp < N2; ++cp) acc = cfma(Sd[cp * N2 + ip], v[cp], acc);
    Zc[(long)ip * ncol] = acc;
  }
}

template <int N2>
__global__ void __launch_bounds__(256) k_gamq_zt(Geom g, int npts, const z64* __restrict__ S, z64* __restrict__ Z) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ncol = (long)npts * N2;
  if (idx >= (long)g.nf * ncol) return;
  const int col = (int)(idx % ncol), d = (int)(idx / ncol);
  const int b = col / N2, s = col - b * N2;
  const z64* Sd = S + ((long)b * g.nf + d) * (N2 * N2) + s * N2;
  z64* Zc = Z + (long)d * N2 * ncol + col;
#pragma unroll
  for (int ip = 0; ip < N2; ++ip) Zc[(long)ip * ncol] = Sd[ip];
}

// k_resid_z / k_resid_zt over both channels: Y (a solution) or, with Y null, (1 + gamma^w)^T
static void resid_z(const Geom& g, int npts, const z64* S, const z64* Y, const z64* opgd, const z64* opgm, z64* Z,
                    hipStream_t s) {
  const long tz = 2L * g.nf * npts * g.n2;
#define ADGA_RZ(NN)                                                                                                 \
  if (g.n == NN) {                                                                                                 \
    hipLaunchKernelGGL(k_resid_zt<NN * NN>, dim3((unsigned)((tz + 255) / 256)), dim3(256), 0, s, g, npts, S, Y,     \
                       opgd, opgm, Z);                                                                             \
    return;                                                                                                        \
  }
  ADGA_RZ(1) ADGA_RZ(2) ADGA_RZ(3)
#undef ADGA_RZ
  const long tot = 2L * npts * g.N * g.n2;
  hipLaunchKernelGGL(k_resid_z, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, S, Y, opgd, opgm, Z);
}

void launch_gamq_z(const Geom& g, int npts, const z64* S, z64* Z, hipStream_t s) {
  const long tz = (long)g.nf * npts * g.n2;
#define ADGA_GZ(NN)                                                                                                 \
  if (g.n == NN) {                                                                                                 \
    hipLaunchKernelGGL(k_gamq_zt<NN * NN>, dim3((unsigned)((tz + 255) / 256)), dim3(256), 0, s, g, npts, S, Z);     \
    return;                                                                                                        \
  }
  ADGA_GZ(1) ADGA_GZ(2) ADGA_GZ(3)
#undef ADGA_GZ
  long tot = (long)npts * g.N * g.n2;
  hipLaunchKernelGGL(k_gamq_z, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, S, Z);
}

// gamma^q_d for every point of the batch, into G2 with the residual layout (point b at 2*b).
void launch_gamq(const Geom& g, int npts, const z64* S, const z64* cd, z64* Z, z64* G2, hipStream_t s) {
  launch_gamq_z(g, npts, S, Z, s);
  resid_gemm(g, npts, 0, 1, cd, cd, Z, G2, nullptr, nullptr, s);
}

// Right-hand side of the eta system: eta A = (1+gamma^w)(1 - A) = (1+gamma^w) blockdiag(S) C, i.e.
// B = M eta^T with B = C^T (blockdiag(S)^T (1+gamma^w)^T), in FP64; with V(q) the density channel adds
// (1+gamma^w_d)^T (W^T (1+gamma^w_d)^T).
void launch_rhs(const Geom& g, int npts, const z64* S, const z64* cd, const z64* cm, const z64* opgd,
                const z64* opgm, z64* Z, z64* B, hipStream_t s, const VqTerm* vq) {
  resid_z(g, npts, S, nullptr, opgd, opgm, Z, s);
  resid_gemm(g, npts, 0, 2, cd, cm, Z, B, nullptr, nullptr, s);
  if (vq) launch_vq_add(g, npts, *vq, nullptr, B, s);
}

// R = B - M Y with M Y = Y - C^T (blockdiag(S)^T Y), in FP64 from the unrounded factors S and C (with V(q):
// minus (1 + gamma^w_d)^T (W^T Y) in the density channel).
void launch_residual(const Geom& g, int npts, const z64* S, const z64* cd, const z64* cm, const z64* B,
                     const z64* Y, z64* Z, z64* R, double* rel, hipStream_t s, const VqTerm* vq) {
  const int nmat = 2 * npts;
  resid_z(g, npts, S, Y, nullptr, nullptr, Z, s);
  resid_gemm(g, npts, 0, 2, cd, cm, Z, R, B, Y, s);
  if (vq) launch_vq_add(g, npts, *vq, Y, R, s);
  hipLaunchKernelGGL(k_resid_norm, dim3(nmat), dim3(256), 0, s, g, R, B, rel);
}

}  // namespace adga
