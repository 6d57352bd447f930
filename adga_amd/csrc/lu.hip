// A7/A8: batched complex LU with partial pivoting and the n^2 right-hand-side solves, one
// workgroup per matrix, on gfx950 MFMA.
//
// M is row-major (row r, column c at r*ldm + c) with the matrix in columns [0, N) and the
// right-hand sides in [N, W).  The factorisation is recursive (left half, Schur update of the
// right half, right half), written out on the host as a list of operations so the kernel is a
// flat loop:
//   LEAF(c0, c1)                     columns [c0, c1) of rows [c0, N) factored in registers with
//                                    partial pivoting (|re|+|im| like izamax, first index on ties);
//                                    its interchanges are applied to whole rows (LAPACK convention),
//                                    which are contiguous in this layout
//   TRSM_L(k0, k1, c0, c1)           rows [k0,k1) of columns [c0,c1) <- L(k0:k1,k0:k1)^-1 (unit lower)
//   TRSM_U(k0, k1, c0, c1)           rows [k0,k1) of columns [c0,c1) <- U(k0:k1,k0:k1)^-1
//   GEMM(r0, r1, c0, c1, k0, k1)     M[r0:r1, c0:c1] -= M[r0:r1, k0:k1] M[k0:k1, c0:c1]
//   SWAP_RHS                         all interchanges, in order, on columns [N, W)
// Every large triangular solve is itself recursive, so nearly all flops run in GEMMs with a
// large K.  The GEMM runs on MFMA (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64): a complex
// product is a real product with K doubled and re/im interleaved along it, the lane's K slot
// selecting (-re, +im) or (-im, -re) of the A element and (re, im) of the B element.
#include <vector>

#include "adga_kernels.h"

namespace adga {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// Operand layout of both 16x16x4 instructions: lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Accumulator: register v of lane l holds D[row(l/16, v)][l%16], with row = 4*(l/16) + v for the
// f32 instruction and row = l/16 + 4*v for the f64 one (both measured on gfx950).
template <>
struct Mfma<float> {
  typedef f32x4 V;
  static __device__ __forceinline__ V op(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lg, int v) { return 4 * lg + v; }
};
template <>
struct Mfma<double> {
  typedef f64x4 V;
  static __device__ __forceinline__ V op(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lg, int v) { return lg + 4 * v; }
};

constexpr int NT = 256;           // threads per workgroup (4 waves)
constexpr int TM = 64, TN = 64;   // workgroup GEMM tile (complex elements); 32x32 per wave
constexpr int KC = 32;            // K chunk staged through LDS

enum { OP_LEAF = 0, OP_TRSM_L = 1, OP_TRSM_U = 2, OP_GEMM = 3, OP_SWAP_RHS = 4 };

template <typename T, int LW>
struct LuLds {
  cplx<T> At[KC][TM + 4];  // A chunk, transposed: At[k][row]
  cplx<T> Bs[KC][TN + 4];  // B chunk: Bs[k][col]
  cplx<T> sP[LW], sQ[LW];
  T redv[NT / 64];
  int redi[NT / 64];
  int spiv[LW];
};

// ---------------------------------------------------------------- GEMM on MFMA
// C[r0:r1, c0:c1] -= A[r0:r1, k0:k1] B[k0:k1, c0:c1] over 64x64 output tiles (32x32 per wave).
// K is staged through LDS in chunks of KC; the next chunk's global loads are issued into
// registers before the current chunk's MFMAs, so load latency overlaps the matrix work.
template <typename T, int LW>
__device__ __forceinline__ void wg_gemm(cplx<T>* __restrict__ M, int ldm, int r0, int r1, int c0, int c1, int k0,
                                        int k1, LuLds<T, LW>& L) {
  using V = typename Mfma<T>::V;
  using C = cplx<T>;
  constexpr int PA = TM * KC / NT, PB = KC * TN / NT;  // prefetched elements per thread
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lg = lane >> 4, part = lg & 1, kcl = lg >> 1;
  for (int rt = r0; rt < r1; rt += TM) {
    for (int ct = c0; ct < c1; ct += TN) {
      V accr[2][2], acci[2][2];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            int row = rt + wr + rb * 16 + Mfma<T>::row(lg, v), col = ct + wc + cb * 16 + li;
            C c = (row < r1 && col < c1) ? M[(long)row * ldm + col] : mk<T>(0, 0);
            accr[rb][cb][v] = c.re;
            acci[rb][cb][v] = c.im;
          }
      C pa[PA], pb[PB];
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        int e = t + i * NT, row = rt + e / KC, k = k0 + e % KC;
        pa[i] = (row < r1 && k < k1) ? M[(long)row * ldm + k] : mk<T>(0, 0);
      }
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        int e = t + i * NT, k = k0 + e / TN, col = ct + e % TN;
        pb[i] = (k < k1 && col < c1) ? M[(long)k * ldm + col] : mk<T>(0, 0);
      }
      for (int kb = k0; kb < k1; kb += KC) {
        __syncthreads();  // the previous chunk's MFMAs are done with LDS
#pragma unroll
        for (int i = 0; i < PA; ++i) {
          int e = t + i * NT;
          L.At[e % KC][e / KC] = pa[i];
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
          int e = t + i * NT;
          L.Bs[e / TN][e % TN] = pb[i];
        }
        __syncthreads();
        const int kn = kb + KC;
        if (kn < k1) {
#pragma unroll
          for (int i = 0; i < PA; ++i) {
            int e = t + i * NT, row = rt + e / KC, k = kn + e % KC;
            pa[i] = (row < r1 && k < k1) ? M[(long)row * ldm + k] : mk<T>(0, 0);
          }
#pragma unroll
          for (int i = 0; i < PB; ++i) {
            int e = t + i * NT, k = kn + e / TN, col = ct + e % TN;
            pb[i] = (k < k1 && col < c1) ? M[(long)k * ldm + col] : mk<T>(0, 0);
          }
        }
        const int kend = min(KC, k1 - kb);
#pragma unroll 4
        for (int kk = 0; kk < kend; kk += 2) {
          T ar[2], ai[2], bb[2];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            C a = L.At[kk + kcl][wr + rb * 16 + li];
            ar[rb] = part ? a.im : -a.re;
            ai[rb] = part ? -a.re : -a.im;
          }
#pragma unroll
          for (int cb = 0; cb < 2; ++cb) {
            C b = L.Bs[kk + kcl][wc + cb * 16 + li];
            bb[cb] = part ? b.im : b.re;
          }
#pragma unroll
          for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
              accr[rb][cb] = Mfma<T>::op(ar[rb], bb[cb], accr[rb][cb]);
              acci[rb][cb] = Mfma<T>::op(ai[rb], bb[cb], acci[rb][cb]);
            }
        }
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            int row = rt + wr + rb * 16 + Mfma<T>::row(lg, v), col = ct + wc + cb * 16 + li;
            if (row < r1 && col < c1) M[(long)row * ldm + col] = mk<T>(accr[rb][cb][v], acci[rb][cb][v]);
          }
      __syncthreads();  // LDS reuse by the next tile
    }
  }
}

// ---------------------------------------------------------------- leaf factorisation
template <typename T, int RPT, int LW>
__device__ __forceinline__ void wg_leaf(cplx<T>* __restrict__ M, int ldm, int N, int W, int c0, int c1,
                                        int* __restrict__ piv, LuLds<T, LW>& L, int& myinfo) {
  using C = cplx<T>;
  constexpr int NWAVE = NT / 64;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int w = c1 - c0;
  C reg[RPT][LW];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    int r = c0 + t + i * NT;
#pragma unroll
    for (int j = 0; j < LW; ++j) reg[i][j] = (r < N && j < w) ? M[(long)r * ldm + c0 + j] : mk<T>(0, 0);
  }
#pragma unroll
  for (int j = 0; j < LW; ++j) {
    if (j < w) {
      const int kc = c0 + j;
      T bv = (T)-1;
      int br = 0x7fffffff;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        int r = c0 + t + i * NT;
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
      if (lane == 0) { L.redv[wave] = bv; L.redi[wave] = br; }
      __syncthreads();
      bv = L.redv[0];
      br = L.redi[0];
#pragma unroll
      for (int q = 1; q < NWAVE; ++q) {
        T ov = L.redv[q];
        int orow = L.redi[q];
        if (ov > bv || (ov == bv && orow < br)) { bv = ov; br = orow; }
      }
      // no finite candidate (NaN column): keep the diagonal row and flag the matrix
      const int p = (br < N) ? br : kc;
      if (br >= N && myinfo == 0) myinfo = kc + 1;
      const int tp = (p - c0) % NT, sp = (p - c0) / NT;
      if (t == tp) {
#pragma unroll
        for (int i = 0; i < RPT; ++i)
          if (i == sp) {
#pragma unroll
            for (int jj = 0; jj < LW; ++jj) L.sP[jj] = reg[i][jj];
          }
      }
      if (p != kc && t == j) {
#pragma unroll
        for (int jj = 0; jj < LW; ++jj) L.sQ[jj] = reg[0][jj];
      }
      if (t == 0) { L.spiv[j] = p; piv[kc] = p; }
      __syncthreads();
      if (p != kc) {
        if (t == tp) {
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (i == sp) {
#pragma unroll
              for (int jj = 0; jj < LW; ++jj) reg[i][jj] = L.sQ[jj];
            }
        }
        if (t == j) {
#pragma unroll
          for (int jj = 0; jj < LW; ++jj) reg[0][jj] = L.sP[jj];
        }
      }
      const C pv = L.sP[j];
      if (pv.re == (T)0 && pv.im == (T)0) {
        if (myinfo == 0) myinfo = kc + 1;
      } else {
        const C rp = crecip(pv);
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          int r = c0 + t + i * NT;
          if (r < N && r > kc) {
            C l = cmul(reg[i][j], rp);
            reg[i][j] = l;
#pragma unroll
            for (int jj = j + 1; jj < LW; ++jj) reg[i][jj] = cfms(l, L.sP[jj], reg[i][jj]);
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    int r = c0 + t + i * NT;
    if (r < N) {
#pragma unroll
      for (int j = 0; j < LW; ++j)
        if (j < w) M[(long)r * ldm + c0 + j] = reg[i][j];
    }
  }
  __syncthreads();
  // interchanges on every other column of the rows (contiguous in the row-major layout)
  bool any = false;
  for (int j = 0; j < w; ++j) any |= (L.spiv[j] != c0 + j);
  if (any) {
    for (int c = t; c < W; c += NT) {
      if (c >= c0 && c < c1) continue;
      for (int j = 0; j < w; ++j) {
        int p = L.spiv[j];
        if (p != c0 + j) {
          long a0 = (long)(c0 + j) * ldm + c, a1 = (long)p * ldm + c;
          C u = M[a0];
          M[a0] = M[a1];
          M[a1] = u;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- small triangular solves
template <typename T, int LW>
__device__ __forceinline__ void wg_trsm_l(cplx<T>* __restrict__ M, int ldm, int k0, int k1, int c0, int c1) {
  using C = cplx<T>;
  const int w = k1 - k0;
  for (int c = c0 + threadIdx.x; c < c1; c += NT) {
    C x[LW];
#pragma unroll
    for (int j = 0; j < LW; ++j)
      if (j < w) x[j] = M[(long)(k0 + j) * ldm + c];
#pragma unroll
    for (int j = 1; j < LW; ++j)
      if (j < w) {
#pragma unroll
        for (int i = 0; i < j; ++i) x[j] = cfms(M[(long)(k0 + j) * ldm + k0 + i], x[i], x[j]);
      }
#pragma unroll
    for (int j = 1; j < LW; ++j)
      if (j < w) M[(long)(k0 + j) * ldm + c] = x[j];
  }
}

template <typename T, int LW>
__device__ __forceinline__ void wg_trsm_u(cplx<T>* __restrict__ M, int ldm, int k0, int k1, int c0, int c1) {
  using C = cplx<T>;
  const int w = k1 - k0;
  for (int c = c0 + threadIdx.x; c < c1; c += NT) {
    C x[LW];
#pragma unroll
    for (int j = 0; j < LW; ++j)
      if (j < w) x[j] = M[(long)(k0 + j) * ldm + c];
#pragma unroll
    for (int j = LW - 1; j >= 0; --j)
      if (j < w) {
#pragma unroll
        for (int i = j + 1; i < LW; ++i)
          if (i < w) x[j] = cfms(M[(long)(k0 + j) * ldm + k0 + i], x[i], x[j]);
        x[j] = cdiv(x[j], M[(long)(k0 + j) * ldm + k0 + j]);
      }
#pragma unroll
    for (int j = 0; j < LW; ++j)
      if (j < w) M[(long)(k0 + j) * ldm + c] = x[j];
  }
}

// ---------------------------------------------------------------- the kernel
template <typename T, int RPT, int LW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2))) k_lu(int N, int W, int ldm, int nops, const int* __restrict__ ops,
                                           cplx<T>* __restrict__ Mall, int* __restrict__ pivall,
                                           int* __restrict__ info) {
  __shared__ LuLds<T, LW> L;
  cplx<T>* M = Mall + (long)blockIdx.x * N * ldm;
  int* piv = pivall + (long)blockIdx.x * N;
  int myinfo = 0;
  for (int o = 0; o < nops; ++o) {
    const int* op = ops + 8 * o;
    const int ty = op[0];
    if (ty == OP_LEAF) {
      wg_leaf<T, RPT, LW>(M, ldm, N, W, op[1], op[2], piv, L, myinfo);
    } else if (ty == OP_TRSM_L) {
      wg_trsm_l<T, LW>(M, ldm, op[1], op[2], op[3], op[4]);
    } else if (ty == OP_TRSM_U) {
      wg_trsm_u<T, LW>(M, ldm, op[1], op[2], op[3], op[4]);
    } else if (ty == OP_GEMM) {
      wg_gemm<T, LW>(M, ldm, op[1], op[2], op[3], op[4], op[5], op[6], L);
    } else {  // OP_SWAP_RHS
      for (int c = N + threadIdx.x; c < W; c += NT)
        for (int j = 0; j < N; ++j) {
          int p = piv[j];
          if (p != j) {
            cplx<T> u = M[(long)j * ldm + c];
            M[(long)j * ldm + c] = M[(long)p * ldm + c];
            M[(long)p * ldm + c] = u;
          }
        }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && info) info[blockIdx.x] = myinfo;
}

// ---------------------------------------------------------------- host: operation lists
namespace {

struct Ops {
  std::vector<int> v;
  void add(int ty, int a, int b, int c, int d, int e = 0, int f = 0) {
    int o[8] = {ty, a, b, c, d, e, f, 0};
    v.insert(v.end(), o, o + 8);
  }
};

int half_split(int w, int lw) {
  int leaves = (w + lw - 1) / lw;
  return ((leaves + 1) / 2) * lw;
}

// L(k0:k1,k0:k1)^-1 applied to rows [k0,k1) of columns [c0,c1)
void rec_trsm_l(Ops& ops, int lw, int k0, int k1, int c0, int c1) {
  if (c1 <= c0) return;
  if (k1 - k0 <= lw) {
    ops.add(OP_TRSM_L, k0, k1, c0, c1);
    return;
  }
  int h = k0 + half_split(k1 - k0, lw);
  rec_trsm_l(ops, lw, k0, h, c0, c1);
  ops.add(OP_GEMM, h, k1, c0, c1, k0, h);
  rec_trsm_l(ops, lw, h, k1, c0, c1);
}

void rec_trsm_u(Ops& ops, int lw, int k0, int k1, int c0, int c1) {
  if (c1 <= c0) return;
  if (k1 - k0 <= lw) {
    ops.add(OP_TRSM_U, k0, k1, c0, c1);
    return;
  }
  int h = k0 + half_split(k1 - k0, lw);
  rec_trsm_u(ops, lw, h, k1, c0, c1);
  ops.add(OP_GEMM, k0, h, c0, c1, h, k1);
  rec_trsm_u(ops, lw, k0, h, c0, c1);
}

// factor columns [c0, c1); `rhs` = the right-hand-side columns [N, W) ride along with this range
void rec_lu(Ops& ops, int lw, int N, int W, int c0, int c1, bool rhs) {
  if (c1 - c0 <= lw) {
    ops.add(OP_LEAF, c0, c1, 0, 0);
    if (rhs && W > N) {
      rec_trsm_l(ops, lw, c0, c1, N, W);
      if (c1 < N) ops.add(OP_GEMM, c1, N, N, W, c0, c1);
    }
    return;
  }
  int h = c0 + half_split(c1 - c0, lw);
  rec_lu(ops, lw, N, W, c0, h, false);
  int hi = rhs ? W : c1;
  rec_trsm_l(ops, lw, c0, h, h, hi);
  ops.add(OP_GEMM, h, N, h, hi, c0, h);
  rec_lu(ops, lw, N, W, h, c1, rhs);
}

}  // namespace

template <typename T, int RPT>
struct LeafW {
  static constexpr int value = sizeof(T) == 4 ? 8 : 4;
};

static int round_rpt(int rpt) {
  static const int opts[] = {1, 2, 3, 4, 5, 6, 8, 12, 16};
  for (int o : opts)
    if (rpt <= o) return o;
  return -1;
}

template <typename T>
int lu_leaf_width(int N) {
  (void)N;
  return sizeof(T) == 4 ? 8 : 4;
}
template int lu_leaf_width<float>(int);
template int lu_leaf_width<double>(int);

// factor + forward/backward solve of the appended right-hand sides
std::vector<int> lu_ops_factor(int N, int W, int lw) {
  Ops ops;
  rec_lu(ops, lw, N, W, 0, N, true);
  if (W > N) rec_trsm_u(ops, lw, 0, N, N, W);
  return ops.v;
}

// solve only, with the factors in place and new right-hand sides in columns [N, W)
std::vector<int> lu_ops_solve(int N, int W, int lw) {
  Ops ops;
  ops.add(OP_SWAP_RHS, 0, 0, 0, 0);
  rec_trsm_l(ops, lw, 0, N, N, W);
  rec_trsm_u(ops, lw, 0, N, N, W);
  return ops.v;
}

template <typename T>
bool launch_lu(int nmat, int N, int W, int ldm, const int* ops, int nops, cplx<T>* M, int* piv, int* info,
               hipStream_t s) {
  int rpt = round_rpt((N + NT - 1) / NT);
  dim3 grid(nmat), blk(NT);
#define ADGA_LU_CASE(R)                                                                                      \
  case R:                                                                                                    \
    hipLaunchKernelGGL((k_lu<T, R, LeafW<T, R>::value>), grid, blk, 0, s, N, W, ldm, nops, ops, M, piv, info); \
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
template bool launch_lu<float>(int, int, int, int, const int*, int, c32*, int*, int*, hipStream_t);
template bool launch_lu<double>(int, int, int, int, const int*, int, z64*, int*, int*, hipStream_t);

// Y[m][r][s] (FP64) = M_m[r][N + s]  (stage 0),  Y += M_m[r][N + s]  (stage 1),
// M_m[r][N + s] = R[m][r][s] rounded to T  (stage 2)
template <typename T>
__global__ void k_rhs_io(int N, int nrhs, int ldm, int nmat, int stage, cplx<T>* __restrict__ M, z64* __restrict__ Y,
                         const z64* __restrict__ R) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)nmat * N * nrhs;
  if (idx >= tot) return;
  int s = (int)(idx % nrhs);
  long mr = idx / nrhs;  // m*N + r
  cplx<T>* e = M + mr * ldm + N + s;
  if (stage == 0)
    Y[idx] = ccast<double>(*e);
  else if (stage == 1)
    Y[idx] = cadd(Y[idx], ccast<double>(*e));
  else
    *e = ccast<T>(R[idx]);
}
template <typename T>
void launch_rhs_io(int nmat, int N, int nrhs, int ldm, int stage, cplx<T>* M, z64* Y, const z64* R, hipStream_t s) {
  long tot = (long)nmat * N * nrhs;
  hipLaunchKernelGGL(k_rhs_io<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, N, nrhs, ldm, nmat, stage, M,
                     Y, R);
}
template void launch_rhs_io<float>(int, int, int, int, int, c32*, z64*, const z64*, hipStream_t);
template void launch_rhs_io<double>(int, int, int, int, int, z64*, z64*, const z64*, hipStream_t);

}  // namespace adga
