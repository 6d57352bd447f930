// A3 (bubble), A6 (BSE build), A7/A8 (batched complex LU + solve, FP64 or
// complex64 with FP64 iterative refinement) on gfx950.
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

namespace adga {

// ---------------------------------------------------------------- A3: bubble
// chi0q[b][d](i1 = i*n + j, i2 = l*n + k') = -beta/nkp sum_K G(K, nu_d)_il G(K - q_b, nu_d - omega)_k'j
// (accumulate_chi0, one_particle_quant_module.f90:274-288; k-sum main.F90:647-657).
// One block per (nu, point); thread = ((i,l), k-slice); n^2 accumulators per thread.
template <int NN>
__global__ void __launch_bounds__(256) k_bubble(Geom g, int iwb, const int* __restrict__ qidx,
                                                const z64* __restrict__ gt, z64* __restrict__ chi0q) {
  constexpr int N2 = NN * NN;
  constexpr int NKS = 256 / N2;
  __shared__ z64 red[256];
  const int d = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const int og = t % N2, ks = t / N2;
  const int i = og / NN, l = og % NN;
  const int q = qidx[b];
  const int iwf = d - g.iwf;
  const z64* g1 = gt + (long)(iwf - g.glo) * g.nkp * N2;
  const z64* g2 = gt + (long)(iwf - iwb - g.glo) * g.nkp * N2;
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
      out[(ii * NN + jj) + (ll * NN + kp) * N2] = cscale(s, f);
    }
    __syncthreads();
  }
}

void launch_bubble(const Geom& g, int iwb, int npts, const int* qidx, const z64* gt, z64* chi0q, hipStream_t s) {
  dim3 grid(g.nf, npts);
  switch (g.n) {
    case 1: hipLaunchKernelGGL(k_bubble<1>, grid, dim3(256), 0, s, g, iwb, qidx, gt, chi0q); break;
    case 2: hipLaunchKernelGGL(k_bubble<2>, grid, dim3(256), 0, s, g, iwb, qidx, gt, chi0q); break;
    case 3: hipLaunchKernelGGL(k_bubble<3>, grid, dim3(256), 0, s, g, iwb, qidx, gt, chi0q); break;
    case 4: hipLaunchKernelGGL(k_bubble<4>, grid, dim3(256), 0, s, g, iwb, qidx, gt, chi0q); break;
    case 5: hipLaunchKernelGGL(k_bubble<5>, grid, dim3(256), 0, s, g, iwb, qidx, gt, chi0q); break;
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
// Row r of M_{b,ch}: M(r, d*n2 + s) = delta - R_d(s, r), R_d(s, r) = sum_i' S_b[d][s][i'] C_ch(d*n2 + i', r);
// M(r, N + s) = B(r, s), the right-hand side of M eta^T = B (launch_rhs).  For the density channel
// gamma^q_d(s, r) = sum_d R_d(s, r) (main.F90:693-700).  One block per (row, channel, point).
template <typename T>
__global__ void __launch_bounds__(256) k_build(Geom g, int ldm, const z64* __restrict__ S, const z64* __restrict__ c_d,
                                               const z64* __restrict__ c_m, const z64* __restrict__ opg_d,
                                               const z64* __restrict__ opg_m, const z64* __restrict__ B,
                                               cplx<T>* __restrict__ M, z64* __restrict__ gamq) {
  extern __shared__ z64 rrow[];  // N values of R for this row (density channel only)
  const int r = blockIdx.x, ch = blockIdx.y, b = blockIdx.z, t = threadIdx.x;
  const int N = g.N, n2 = g.n2;
  const z64* C = ch == 0 ? c_d : c_m;
  const z64* Sb = S + (long)b * g.nf * g.n4;
  const z64* crow = C + (long)r * N;  // C(., r) is contiguous (column r of the Fortran array)
  cplx<T>* mrow = M + ((long)(2 * b + ch) * N + r) * ldm;
  for (int c = t; c < ldm; c += 256) {
    if (c < N) {
      int d = c / n2, s = c % n2;
      const z64* srow = Sb + (long)d * g.n4 + s * n2;
      const z64* cc = crow + d * n2;
      z64 acc = mk(0.0, 0.0);
      for (int ip = 0; ip < n2; ++ip) acc = cfma(srow[ip], cc[ip], acc);
      if (ch == 0) rrow[c] = acc;
      z64 m = mk((c == r ? 1.0 : 0.0) - acc.re, -acc.im);
      mrow[c] = ccast<T>(m);
    } else if (c < N + n2) {
      mrow[c] = ccast<T>(B[((long)(2 * b + ch) * N + r) * n2 + (c - N)]);
    } else {
      mrow[c] = mk<T>(0, 0);
    }
  }
  if (ch == 0) {
    __syncthreads();
    if (t < n2) {
      z64 acc = mk(0.0, 0.0);
      for (int d = 0; d < g.nf; ++d) acc = cadd(acc, rrow[d * n2 + t]);
      gamq[(long)b * n2 * N + t + (long)r * n2] = acc;
    }
  }
}

template <typename T>
void launch_build(const Geom& g, int npts, int ldm, const z64* S, const z64* cd, const z64* cm, const z64* opgd,
                  const z64* opgm, const z64* B, cplx<T>* M, z64* gamq, hipStream_t s) {
  dim3 grid(g.N, 2, npts);
  size_t lds = (size_t)g.N * sizeof(z64);
  hipLaunchKernelGGL(k_build<T>, grid, dim3(256), lds, s, g, ldm, S, cd, cm, opgd, opgm, B, M, gamq);
}
template void launch_build<float>(const Geom&, int, int, const z64*, const z64*, const z64*, const z64*, const z64*,
                                  const z64*, c32*, z64*, hipStream_t);
template void launch_build<double>(const Geom&, int, int, const z64*, const z64*, const z64*, const z64*, const z64*,
                                   const z64*, z64*, z64*, hipStream_t);

// ---------------------------------------------------------------- A7: batched LU + solve
// One workgroup (256 threads) per matrix.  For each leaf of LW columns:
//   1. the leaf (rows k0..N-1) is held in registers (RPT rows per thread) and factored column by
//      column with partial pivoting (|re|+|im| like izamax, first index on ties);
//   2. the leaf's row interchanges are applied to columns right of the leaf, and the leaf's rows of
//      those columns are solved with the unit-lower L11 (one thread per column);
//   3. the trailing matrix (including the right-hand-side columns) is updated, C -= L21 U12.
// Columns left of the leaf are not permuted: L is kept in elimination order, which is the order
// the refinement solve applies it in.  Finally U Y = Z is solved for the right-hand sides.
constexpr int LU_NT = 256;

template <typename T, int RPT, int LW>
__global__ void __launch_bounds__(LU_NT) k_lu(int N, int W, int ldm, cplx<T>* __restrict__ Mall,
                                              int* __restrict__ pivall, int* __restrict__ info) {
  using C = cplx<T>;
  constexpr int NWAVE = LU_NT / 64;
  constexpr int TS = 64;  // GEMM tile
  __shared__ C sP[LW], sQ[LW];
  __shared__ C sL[LW][LW + 1];
  __shared__ T redv[NWAVE];
  __shared__ int redi[NWAVE];
  __shared__ int spiv[LW];
  __shared__ C tA[TS][LW + 1];
  __shared__ C tB[LW][TS + 1];
  const int t = threadIdx.x;
  const int lane = t % 64, wave = t / 64;
  C* M = Mall + (long)blockIdx.x * N * ldm;
  int* piv = pivall + (long)blockIdx.x * N;
  int myinfo = 0;

  for (int k0 = 0; k0 < N; k0 += LW) {
    const int w = min(LW, N - k0);
    C reg[RPT][LW];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      int r = k0 + t + i * LU_NT;
#pragma unroll
      for (int j = 0; j < LW; ++j) reg[i][j] = (r < N && j < w) ? M[(long)r * ldm + k0 + j] : mk<T>(0, 0);
    }
#pragma unroll
    for (int j = 0; j < LW; ++j) {
      if (j < w) {
        const int kc = k0 + j;
        // ---- pivot search
        T bv = (T)-1;
        int br = 0x7fffffff;
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          int r = k0 + t + i * LU_NT;
          if (r < N && r >= kc) {
            T v = cabs1(reg[i][j]);
            if (v > bv) { bv = v; br = r; }
          }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          T ov = __shfl_xor(bv, off);
          int orow = __shfl_xor(br, off);
          if (ov > bv || (ov == bv && orow < br)) { bv = ov; br = orow; }
        }
        if (lane == 0) { redv[wave] = bv; redi[wave] = br; }
        __syncthreads();
        bv = redv[0];
        br = redi[0];
#pragma unroll
        for (int q = 1; q < NWAVE; ++q) {
          T ov = redv[q];
          int orow = redi[q];
          if (ov > bv || (ov == bv && orow < br)) { bv = ov; br = orow; }
        }
        // no finite candidate (NaN column): keep the diagonal row and flag the matrix
        const int p = (br < N) ? br : kc;
        if (br >= N && myinfo == 0) myinfo = kc + 1;
        const int tp = (p - k0) % LU_NT, sp = (p - k0) / LU_NT;
        // ---- row interchange inside the leaf through LDS
        if (t == tp) {
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (i == sp) {
#pragma unroll
              for (int jj = 0; jj < LW; ++jj) sP[jj] = reg[i][jj];
            }
        }
        if (p != kc && t == j) {
#pragma unroll
          for (int jj = 0; jj < LW; ++jj) sQ[jj] = reg[0][jj];
        }
        if (t == 0) { spiv[j] = p; piv[kc] = p; }
        __syncthreads();
        if (p != kc) {
          if (t == tp) {
#pragma unroll
            for (int i = 0; i < RPT; ++i)
              if (i == sp) {
#pragma unroll
                for (int jj = 0; jj < LW; ++jj) reg[i][jj] = sQ[jj];
              }
          }
          if (t == j) {
#pragma unroll
            for (int jj = 0; jj < LW; ++jj) reg[0][jj] = sP[jj];
          }
        }
        // ---- scale and rank-1 update of the rows below
        const C pv = sP[j];
        if (pv.re == (T)0 && pv.im == (T)0) {
          if (myinfo == 0) myinfo = kc + 1;
        } else {
          const C rp = crecip(pv);
#pragma unroll
          for (int i = 0; i < RPT; ++i) {
            int r = k0 + t + i * LU_NT;
            if (r < N && r > kc) {
              C l = cmul(reg[i][j], rp);
              reg[i][j] = l;
#pragma unroll
              for (int jj = j + 1; jj < LW; ++jj) reg[i][jj] = cfms(l, sP[jj], reg[i][jj]);
            }
          }
        }
      }
    }
    // ---- store the leaf, keep L11 in LDS
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      int r = k0 + t + i * LU_NT;
      if (r < N) {
#pragma unroll
        for (int j = 0; j < LW; ++j)
          if (j < w) M[(long)r * ldm + k0 + j] = reg[i][j];
      }
    }
    if (t < w) {
#pragma unroll
      for (int j = 0; j < LW; ++j) sL[t][j] = reg[0][j];
    }
    __syncthreads();
    // ---- interchanges and U12 rows for the columns right of the leaf
    for (int c = k0 + w + t; c < W; c += LU_NT) {
      C x[LW];
#pragma unroll
      for (int j = 0; j < LW; ++j) {
        if (j < w) {
          int p = spiv[j];
          long a0 = (long)(k0 + j) * ldm + c;
          if (p != k0 + j) {
            long a1 = (long)p * ldm + c;
            C u = M[a0], v = M[a1];
            M[a1] = u;
            x[j] = v;
          } else {
            x[j] = M[a0];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < LW; ++j) {
        if (j < w) {
#pragma unroll
          for (int i = 0; i < j; ++i) x[j] = cfms(sL[j][i], x[i], x[j]);
          M[(long)(k0 + j) * ldm + c] = x[j];
        }
      }
    }
    __syncthreads();
    // ---- trailing update C -= L21 U12 over 64x64 tiles (4x4 complex outputs per thread)
    const int r0 = k0 + w, c0 = k0 + w;
    const int ty = t / 16, tx = t % 16;
    for (int rt = r0; rt < N; rt += TS) {
      for (int ct = c0; ct < W; ct += TS) {
        for (int e = t; e < TS * LW; e += LU_NT) {
          int rr = e / LW, j = e % LW;
          int r = rt + rr;
          tA[rr][j] = (r < N && j < w) ? M[(long)r * ldm + k0 + j] : mk<T>(0, 0);
        }
        for (int e = t; e < TS * LW; e += LU_NT) {
          int j = e / TS, cc = e % TS;
          int c = ct + cc;
          tB[j][cc] = (c < W && j < w) ? M[(long)(k0 + j) * ldm + c] : mk<T>(0, 0);
        }
        __syncthreads();
        C acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int bq = 0; bq < 4; ++bq) {
            int r = rt + ty * 4 + a, c = ct + tx * 4 + bq;
            acc[a][bq] = (r < N && c < W) ? M[(long)r * ldm + c] : mk<T>(0, 0);
          }
#pragma unroll
        for (int j = 0; j < LW; ++j) {
          C av[4], bvv[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) av[a] = tA[ty * 4 + a][j];
#pragma unroll
          for (int bq = 0; bq < 4; ++bq) bvv[bq] = tB[j][tx * 4 + bq];
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int bq = 0; bq < 4; ++bq) acc[a][bq] = cfms(av[a], bvv[bq], acc[a][bq]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int bq = 0; bq < 4; ++bq) {
            int r = rt + ty * 4 + a, c = ct + tx * 4 + bq;
            if (r < N && c < W) M[(long)r * ldm + c] = acc[a][bq];
          }
        __syncthreads();
      }
    }
  }
  // ---- back substitution U Y = Z for the right-hand-side columns [N, W)
  const int nrhs = W - N;
  constexpr int BS = 32;
  for (int i1 = N; i1 > 0; i1 -= BS) {
    const int i0 = max(0, i1 - BS);
    for (int pq = t; pq < (i1 - i0) * nrhs; pq += LU_NT) {
      int i = i0 + pq / nrhs, s = pq % nrhs;
      C acc = M[(long)i * ldm + N + s];
      for (int j = i1; j < N; ++j) acc = cfms(M[(long)i * ldm + j], M[(long)j * ldm + N + s], acc);
      M[(long)i * ldm + N + s] = acc;
    }
    __syncthreads();
    for (int s = t; s < nrhs; s += LU_NT) {
      for (int i = i1 - 1; i >= i0; --i) {
        C acc = M[(long)i * ldm + N + s];
        for (int j = i + 1; j < i1; ++j) acc = cfms(M[(long)i * ldm + j], M[(long)j * ldm + N + s], acc);
        M[(long)i * ldm + N + s] = cdiv(acc, M[(long)i * ldm + i]);
      }
    }
    __syncthreads();
  }
  if (t == 0) info[blockIdx.x] = myinfo;
}

// Solve M x = rhs with the factors of k_lu (same leaf width LW): forward elimination leaf by
// leaf (interchanges, L11, L21), then back substitution with U.  x is overwritten in place
// (row-major N x nrhs, type T).  One workgroup per matrix.
template <typename T, int LW>
__global__ void __launch_bounds__(LU_NT) k_lu_apply(int N, int nrhs, int ldm, const cplx<T>* __restrict__ Mall,
                                                    const int* __restrict__ pivall, cplx<T>* __restrict__ xall) {
  using C = cplx<T>;
  const int t = threadIdx.x;
  const C* M = Mall + (long)blockIdx.x * N * ldm;
  const int* piv = pivall + (long)blockIdx.x * N;
  C* x = xall + (long)blockIdx.x * N * nrhs;
  for (int k0 = 0; k0 < N; k0 += LW) {
    const int w = min(LW, N - k0);
    for (int s = t; s < nrhs; s += LU_NT) {
      for (int j = 0; j < w; ++j) {
        int p = piv[k0 + j];
        if (p != k0 + j) {
          C u = x[(long)(k0 + j) * nrhs + s];
          x[(long)(k0 + j) * nrhs + s] = x[(long)p * nrhs + s];
          x[(long)p * nrhs + s] = u;
        }
      }
      for (int j = 1; j < w; ++j) {
        C acc = x[(long)(k0 + j) * nrhs + s];
        for (int i = 0; i < j; ++i) acc = cfms(M[(long)(k0 + j) * ldm + k0 + i], x[(long)(k0 + i) * nrhs + s], acc);
        x[(long)(k0 + j) * nrhs + s] = acc;
      }
    }
    __syncthreads();
    for (int pq = t; pq < (N - k0 - w) * nrhs; pq += LU_NT) {
      int r = k0 + w + pq / nrhs, s = pq % nrhs;
      C acc = x[(long)r * nrhs + s];
      for (int j = 0; j < w; ++j) acc = cfms(M[(long)r * ldm + k0 + j], x[(long)(k0 + j) * nrhs + s], acc);
      x[(long)r * nrhs + s] = acc;
    }
    __syncthreads();
  }
  constexpr int BS = 32;
  for (int i1 = N; i1 > 0; i1 -= BS) {
    const int i0 = max(0, i1 - BS);
    for (int pq = t; pq < (i1 - i0) * nrhs; pq += LU_NT) {
      int i = i0 + pq / nrhs, s = pq % nrhs;
      C acc = x[(long)i * nrhs + s];
      for (int j = i1; j < N; ++j) acc = cfms(M[(long)i * ldm + j], x[(long)j * nrhs + s], acc);
      x[(long)i * nrhs + s] = acc;
    }
    __syncthreads();
    for (int s = t; s < nrhs; s += LU_NT) {
      for (int i = i1 - 1; i >= i0; --i) {
        C acc = x[(long)i * nrhs + s];
        for (int j = i + 1; j < i1; ++j) acc = cfms(M[(long)i * ldm + j], x[(long)j * nrhs + s], acc);
        x[(long)i * nrhs + s] = cdiv(acc, M[(long)i * ldm + i]);
      }
    }
    __syncthreads();
  }
}

// Leaf width and rows-per-thread chosen so that a leaf fits in registers.
template <typename T, int RPT>
struct LeafWidth {
  static constexpr int value = sizeof(T) == 4 ? (RPT <= 5 ? 8 : (RPT <= 8 ? 8 : 4)) : (RPT <= 5 ? 8 : (RPT <= 8 ? 4 : 2));
};

static int round_rpt(int rpt) {
  static const int opts[] = {1, 2, 3, 4, 5, 6, 8, 12, 16};
  for (int o : opts)
    if (rpt <= o) return o;
  return -1;
}

template <typename T>
int lu_leaf_width(int N) {
  int rpt = round_rpt((N + LU_NT - 1) / LU_NT);
  if (sizeof(T) == 4) return rpt <= 8 ? 8 : 4;
  return rpt <= 5 ? 8 : (rpt <= 8 ? 4 : 2);
}
template int lu_leaf_width<float>(int);
template int lu_leaf_width<double>(int);

template <typename T>
bool launch_lu(int nmat, int N, int W, int ldm, cplx<T>* M, int* piv, int* info, hipStream_t s) {
  int rpt = round_rpt((N + LU_NT - 1) / LU_NT);
  dim3 grid(nmat), blk(LU_NT);
#define ADGA_LU_CASE(R)                                                                                         \
  case R:                                                                                                       \
    hipLaunchKernelGGL((k_lu<T, R, LeafWidth<T, R>::value>), grid, blk, 0, s, N, W, ldm, M, piv, info); \
    return true;
  switch (rpt) {
    ADGA_LU_CASE(1)
    ADGA_LU_CASE(2)
    ADGA_LU_CASE(3)
    ADGA_LU_CASE(4)
    ADGA_LU_CASE(5)
    ADGA_LU_CASE(6)
    ADGA_LU_CASE(8)
    ADGA_LU_CASE(12)
    ADGA_LU_CASE(16)
    default: return false;
  }
#undef ADGA_LU_CASE
}
template bool launch_lu<float>(int, int, int, int, c32*, int*, int*, hipStream_t);
template bool launch_lu<double>(int, int, int, int, z64*, int*, int*, hipStream_t);

template <typename T>
bool launch_lu_apply(int nmat, int N, int nrhs, int ldm, const cplx<T>* M, const int* piv, cplx<T>* x,
                     hipStream_t s) {
  int lw = lu_leaf_width<T>(N);
  dim3 grid(nmat), blk(LU_NT);
  switch (lw) {
    case 16: hipLaunchKernelGGL((k_lu_apply<T, 16>), grid, blk, 0, s, N, nrhs, ldm, M, piv, x); return true;
    case 8: hipLaunchKernelGGL((k_lu_apply<T, 8>), grid, blk, 0, s, N, nrhs, ldm, M, piv, x); return true;
    case 4: hipLaunchKernelGGL((k_lu_apply<T, 4>), grid, blk, 0, s, N, nrhs, ldm, M, piv, x); return true;
    case 2: hipLaunchKernelGGL((k_lu_apply<T, 2>), grid, blk, 0, s, N, nrhs, ldm, M, piv, x); return true;
    default: return false;
  }
}
template bool launch_lu_apply<float>(int, int, int, int, const c32*, const int*, c32*, hipStream_t);
template bool launch_lu_apply<double>(int, int, int, int, const z64*, const int*, z64*, hipStream_t);

// ---------------------------------------------------------------- solution extraction / refinement
// Y[m][r][s] (FP64) = M_m[r][N + s]
template <typename T>
__global__ void k_extract(int N, int nrhs, int ldm, int nmat, const cplx<T>* __restrict__ M, z64* __restrict__ Y) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)nmat * N * nrhs;
  if (idx >= tot) return;
  int s = (int)(idx % nrhs);
  long mr = idx / nrhs;  // m*N + r
  Y[idx] = ccast<double>(M[mr * ldm + N + s]);
}
template <typename T>
void launch_extract(int nmat, int N, int nrhs, int ldm, const cplx<T>* M, z64* Y, hipStream_t s) {
  long tot = (long)nmat * N * nrhs;
  hipLaunchKernelGGL(k_extract<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, N, nrhs, ldm, nmat, M, Y);
}
template void launch_extract<float>(int, int, int, int, const c32*, z64*, hipStream_t);
template void launch_extract<double>(int, int, int, int, const z64*, z64*, hipStream_t);

// Z[m][d*n2 + i'][s] = sum_c' S_b[d][c'][i'] V[m][d*n2 + c'][s]   (m = 2b + ch), i.e. Z = blockdiag(S)^T V,
// with V = Y (a solution) or, when Y is null, V = (1 + gamma^w_ch)^T.
__global__ void k_resid_z(Geom g, int nmat, const z64* __restrict__ S, const z64* __restrict__ Y,
                          const z64* __restrict__ opg_d, const z64* __restrict__ opg_m, z64* __restrict__ Z) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)nmat * g.N * g.n2;
  if (idx >= tot) return;
  int s = (int)(idx % g.n2);
  long mx = idx / g.n2;
  int x = (int)(mx % g.N);
  int m = (int)(mx / g.N);
  int b = m / 2;
  int d = x / g.n2, ip = x % g.n2;
  const z64* Sd = S + ((long)b * g.nf + d) * g.n4;
  z64 acc = mk(0.0, 0.0);
  if (Y) {
    const z64* Ym = Y + (long)m * g.N * g.n2;
    for (int cp = 0; cp < g.n2; ++cp) acc = cfma(Sd[cp * g.n2 + ip], Ym[(long)(d * g.n2 + cp) * g.n2 + s], acc);
  } else {
    const z64* opg = (m % 2 == 0) ? opg_d : opg_m;
    for (int cp = 0; cp < g.n2; ++cp) acc = cfma(Sd[cp * g.n2 + ip], opg[s + (long)(d * g.n2 + cp) * g.n2], acc);
  }
  Z[idx] = acc;
}

// R[m][r][s] = B[m][r][s] - Y[m][r][s] + sum_x C_ch(x, r) Z[m][x][s]  (B, Y may be null = 0);
// one channel per launch.  GEMM over the shared C: rows r, columns (point, s); 64x64 tiles,
// 16-deep K chunks.
__global__ void __launch_bounds__(256) k_resid_gemm(Geom g, int npts, int ch, const z64* __restrict__ C,
                                                    const z64* __restrict__ B, const z64* __restrict__ Y,
                                                    const z64* __restrict__ Z, z64* __restrict__ R) {
  constexpr int TS = 64, KC = 16;
  __shared__ z64 tA[TS][KC + 1];
  __shared__ z64 tB[KC][TS + 1];
  const int N = g.N, n2 = g.n2;
  const int ncol = npts * n2;
  const int rt = blockIdx.x * TS, ct = blockIdx.y * TS;
  const int t = threadIdx.x, ty = t / 16, tx = t % 16;
  z64 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = mk(0.0, 0.0);
  for (int x0 = 0; x0 < N; x0 += KC) {
    for (int e = t; e < TS * KC; e += 256) {
      int rr = e / KC, kk = e % KC;
      int r = rt + rr, x = x0 + kk;
      tA[rr][kk] = (r < N && x < N) ? C[(long)x + (long)r * N] : mk(0.0, 0.0);  // C(x, r)
    }
    for (int e = t; e < TS * KC; e += 256) {
      int kk = e / TS, cc = e % TS;
      int col = ct + cc, x = x0 + kk;
      z64 v = mk(0.0, 0.0);
      if (col < ncol && x < N) {
        int b = col / n2, s = col % n2;
        v = Z[((long)(2 * b + ch) * N + x) * n2 + s];
      }
      tB[kk][cc] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
      z64 av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = tA[ty * 4 + a][kk];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = tB[kk][tx * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = cfma(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int r = rt + ty * 4 + a, col = ct + tx * 4 + b;
      if (r < N && col < ncol) {
        int pb = col / n2, s = col % n2;
        long yi = ((long)(2 * pb + ch) * N + r) * n2 + s;
        z64 v = B ? B[yi] : mk(0.0, 0.0);
        if (Y) v = csub(v, Y[yi]);
        R[yi] = cadd(v, acc[a][b]);
      }
    }
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

// x (type T) = R for matrices that are still refining; Y += x afterwards.
template <typename T>
__global__ void k_round(long tot, const z64* __restrict__ R, cplx<T>* __restrict__ x) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < tot) x[idx] = ccast<T>(R[idx]);
}
template <typename T>
__global__ void k_accum(long tot, const cplx<T>* __restrict__ x, z64* __restrict__ Y) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < tot) Y[idx] = cadd(Y[idx], ccast<double>(x[idx]));
}

// Right-hand side of the eta system: eta A = (1+gamma^w)(1 - A) = (1+gamma^w) blockdiag(S) C, i.e.
// B = M eta^T with B = C^T (blockdiag(S)^T (1+gamma^w)^T), in FP64.
void launch_rhs(const Geom& g, int npts, const z64* S, const z64* cd, const z64* cm, const z64* opgd,
                const z64* opgm, z64* Z, z64* B, hipStream_t s) {
  int nmat = 2 * npts;
  long tot = (long)nmat * g.N * g.n2;
  hipLaunchKernelGGL(k_resid_z, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, nmat, S, nullptr, opgd,
                     opgm, Z);
  dim3 grid((g.N + 63) / 64, (npts * g.n2 + 63) / 64);
  hipLaunchKernelGGL(k_resid_gemm, grid, dim3(256), 0, s, g, npts, 0, cd, nullptr, nullptr, Z, B);
  hipLaunchKernelGGL(k_resid_gemm, grid, dim3(256), 0, s, g, npts, 1, cm, nullptr, nullptr, Z, B);
}

// R = B - M Y with M Y = Y - C^T (blockdiag(S)^T Y), in FP64 from the unrounded factors S and C.
void launch_residual(const Geom& g, int npts, const z64* S, const z64* cd, const z64* cm, const z64* B,
                     const z64* Y, z64* Z, z64* R, double* rel, hipStream_t s) {
  int nmat = 2 * npts;
  long tot = (long)nmat * g.N * g.n2;
  hipLaunchKernelGGL(k_resid_z, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, nmat, S, Y, nullptr,
                     nullptr, Z);
  dim3 grid((g.N + 63) / 64, (npts * g.n2 + 63) / 64);
  hipLaunchKernelGGL(k_resid_gemm, grid, dim3(256), 0, s, g, npts, 0, cd, B, Y, Z, R);
  hipLaunchKernelGGL(k_resid_gemm, grid, dim3(256), 0, s, g, npts, 1, cm, B, Y, Z, R);
  hipLaunchKernelGGL(k_resid_norm, dim3(nmat), dim3(256), 0, s, g, R, B, rel);
}

template <typename T>
void launch_correct(long tot, const z64* R, cplx<T>* x, int stage, z64* Y, hipStream_t s) {
  dim3 grid((unsigned)((tot + 255) / 256));
  if (stage == 0)
    hipLaunchKernelGGL(k_round<T>, grid, dim3(256), 0, s, tot, R, x);
  else
    hipLaunchKernelGGL(k_accum<T>, grid, dim3(256), 0, s, tot, x, Y);
}
template void launch_correct<float>(long, const z64*, c32*, int, z64*, hipStream_t);
template void launch_correct<double>(long, const z64*, z64*, int, z64*, hipStream_t);

}  // namespace adga
