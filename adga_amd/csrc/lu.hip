// A7/A8: batched complex LU with partial pivoting and the n^2 right-hand-side solves, one
// workgroup per matrix, on gfx950 MFMA.
//
// M is row-major (row r, column c at r*ldm + c) with the matrix in columns [0, N) and the
// right-hand sides in [N, W).  The factorisation is recursive (left half, Schur update of the
// right half, right half) down to panels of P columns, written out on the host as a list of
// operations so the kernel is a flat loop:
//   PANEL(c0, c1)                    columns [c0, c1) (<= P) of rows [c0, N), factored left-looking
//                                    leaf by leaf (LW columns in registers) with partial pivoting
//                                    (|re|+|im| like izamax, first index on ties); each leaf's
//                                    interchanges are applied to whole rows (LAPACK convention),
//                                    which are contiguous in this layout; then the inverses of the
//                                    panel's unit-lower L11 and upper U11 go to a side buffer
//   TRSM_L(k0, k1, c0, c1)           rows [k0,k1) of columns [c0,c1) <- L11^-1 (.) : one GEMM with the
//                                    stored inverse of panel k0
//   TRSM_U(k0, k1, c0, c1)           rows [k0,k1) of columns [c0,c1) <- U11^-1 (.)
//   GEMM(r0, r1, c0, c1, k0, k1)     M[r0:r1, c0:c1] -= M[r0:r1, k0:k1] M[k0:k1, c0:c1]
//   SWAP_RHS                         all interchanges, in order, on columns [N, W)
// Large triangular solves are recursive too, so the work runs as GEMMs.  The GEMM runs on MFMA
// (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64): a complex product is a real product with K
// doubled and re/im interleaved along it, the lane's K slot selecting (re, -im) or (im, re) of the
// A element and (re, im) of the B element.
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
constexpr int P = 32;             // panel width: the recursion's leaf, and the TRSM block
static_assert(P <= TM, "a TRSM block is one row tile (the in-place solve relies on it)");

// K chunk staged through LDS: 32 for complex64, 16 for complex128 (keeps two workgroups per CU)
template <typename T>
struct Kc {
  static constexpr int value = sizeof(T) == 4 ? 32 : 16;
};

enum { OP_PANEL = 0, OP_TRSM_L = 1, OP_TRSM_U = 2, OP_GEMM = 3, OP_SWAP_RHS = 4 };

template <typename T, int LW>
struct LuLds {
  static constexpr int KC = Kc<T>::value;
  union {
    struct {
      cplx<T> At[KC][TM + 4];  // A chunk, transposed: At[k][row]
      cplx<T> Bs[KC][TN + 4];  // B chunk: Bs[k][col]
    } g;
    struct {
      cplx<T> Lb[P][P + 1];     // a panel's diagonal block
      cplx<T> Xb[2][P][P + 1];  // its inverses being formed (L, U)
    } v;
  } u;
  cplx<T> sP[LW], sQ[LW];          // pivot row exchange
  cplx<T> Xs[2][LW][LW + 1];       // left-looking: a previous leaf's rows of the current leaf
  cplx<T> Li[P / LW][LW][LW + 1];  // unit-lower inverses of the panel's leaves
  T redv[NT / 64];
  int redi[NT / 64];
  int spiv[LW];
};

// ---------------------------------------------------------------- GEMM on MFMA
// C[0:m, 0:n] (ldc) <- C - A B (assign = false) or A B (assign = true), A[0:m, 0:K] (lda),
// B[0:K, 0:n] (ldb); 64x64 output tiles (32x32 per wave).  K is staged through LDS in chunks of
// KC; the next chunk's global loads are issued into registers before the current chunk's MFMAs,
// so load latency overlaps the matrix work.  A tile's output is written after all of its K has been
// read by every wave, so C may alias B when one row tile spans all K rows of B (in-place solves).
template <typename T, int LW>
__device__ __forceinline__ void wg_gemm(const cplx<T>* A, int lda, const cplx<T>* B, int ldb, cplx<T>* Cm, int ldc,
                                        int m, int n, int K, bool assign, LuLds<T, LW>& L) {
  using V = typename Mfma<T>::V;
  using C = cplx<T>;
  constexpr int KC = Kc<T>::value;
  constexpr int PA = TM * KC / NT, PB = KC * TN / NT;  // prefetched elements per thread
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int li = lane & 15, lg = lane >> 4, part = lg & 1, kcl = lg >> 1;
  const T sgn = assign ? (T)1 : (T)-1;
  for (int rt = 0; rt < m; rt += TM) {
    for (int ct = 0; ct < n; ct += TN) {
      V accr[2][2], acci[2][2];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            int row = rt + wr + rb * 16 + Mfma<T>::row(lg, v), col = ct + wc + cb * 16 + li;
            C c = (!assign && row < m && col < n) ? Cm[(long)row * ldc + col] : mk<T>(0, 0);
            accr[rb][cb][v] = c.re;
            acci[rb][cb][v] = c.im;
          }
      C pa[PA], pb[PB];
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        int e = t + i * NT, row = rt + e / KC, k = e % KC;
        pa[i] = (row < m && k < K) ? A[(long)row * lda + k] : mk<T>(0, 0);
      }
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        int e = t + i * NT, k = e / TN, col = ct + e % TN;
        pb[i] = (k < K && col < n) ? B[(long)k * ldb + col] : mk<T>(0, 0);
      }
      for (int kb = 0; kb < K; kb += KC) {
        __syncthreads();  // the previous chunk's MFMAs are done with LDS
#pragma unroll
        for (int i = 0; i < PA; ++i) {
          int e = t + i * NT;
          L.u.g.At[e % KC][e / KC] = pa[i];
        }
#pragma unroll
        for (int i = 0; i < PB; ++i) {
          int e = t + i * NT;
          L.u.g.Bs[e / TN][e % TN] = pb[i];
        }
        __syncthreads();
        const int kn = kb + KC;
        if (kn < K) {
#pragma unroll
          for (int i = 0; i < PA; ++i) {
            int e = t + i * NT, row = rt + e / KC, k = kn + e % KC;
            pa[i] = (row < m && k < K) ? A[(long)row * lda + k] : mk<T>(0, 0);
          }
#pragma unroll
          for (int i = 0; i < PB; ++i) {
            int e = t + i * NT, k = kn + e / TN, col = ct + e % TN;
            pb[i] = (k < K && col < n) ? B[(long)k * ldb + col] : mk<T>(0, 0);
          }
        }
        const int kend = min(KC, K - kb);
#pragma unroll 4
        for (int kk = 0; kk < kend; kk += 2) {
          T ar[2], ai[2], bb[2];
#pragma unroll
          for (int rb = 0; rb < 2; ++rb) {
            C a = L.u.g.At[kk + kcl][wr + rb * 16 + li];
            ar[rb] = sgn * (part ? -a.im : a.re);
            ai[rb] = sgn * (part ? a.re : a.im);
          }
#pragma unroll
          for (int cb = 0; cb < 2; ++cb) {
            C b = L.u.g.Bs[kk + kcl][wc + cb * 16 + li];
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
      __syncthreads();  // every wave has consumed its operands before a tile output (may alias B) lands
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int cb = 0; cb < 2; ++cb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            int row = rt + wr + rb * 16 + Mfma<T>::row(lg, v), col = ct + wc + cb * 16 + li;
            if (row < m && col < n) Cm[(long)row * ldc + col] = mk<T>(accr[rb][cb][v], acci[rb][cb][v]);
          }
    }
  }
}

// ---------------------------------------------------------------- panel factorisation
// Columns [c0, c1) of rows [c0, N); thread t holds rows c0 + t + i*NT (i < RPT) of the current leaf.
template <typename T, int RPT, int LW>
__device__ __forceinline__ void wg_panel(cplx<T>* __restrict__ M, int ldm, int N, int W, int c0, int c1,
                                         int* __restrict__ piv, cplx<T>* __restrict__ inv, LuLds<T, LW>& L,
                                         int& myinfo) {
  using C = cplx<T>;
  constexpr int NWAVE = NT / 64;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int nl = (c1 - c0 + LW - 1) / LW;
  int xb = 0;  // double-buffer index of Xs
  for (int jl = 0; jl < nl; ++jl) {
    const int j0 = c0 + jl * LW;
    const int w = min(LW, c1 - j0);
    C reg[RPT][LW];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      int r = c0 + t + i * NT;
#pragma unroll
      for (int j = 0; j < LW; ++j) reg[i][j] = (r < N && j < w) ? M[(long)r * ldm + j0 + j] : mk<T>(0, 0);
    }
    // ---- left-looking updates from the panel's previous leaves
    for (int il = 0; il < jl; ++il) {
      const int p0 = c0 + il * LW;
      const int a = t - (p0 - c0);  // this thread owns row p0 + a (slot 0) when 0 <= a < LW
      const bool owner = a >= 0 && a < LW;
      if (owner) {
#pragma unroll
        for (int j = 0; j < LW; ++j) L.Xs[xb][a][j] = reg[0][j];
      }
      __syncthreads();
      C xn[LW];
      if (owner) {  // rows of the U block: L_ii^-1 applied
#pragma unroll
        for (int j = 0; j < LW; ++j) {
          C acc = mk<T>(0, 0);
#pragma unroll
          for (int b = 0; b < LW; ++b)
            if (b <= a) acc = cfma(L.Li[il][a][b], L.Xs[xb][b][j], acc);
          xn[j] = acc;
        }
      }
      __syncthreads();
      if (owner) {
#pragma unroll
        for (int j = 0; j < LW; ++j) {
          reg[0][j] = xn[j];
          L.Xs[xb][a][j] = xn[j];
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        int r = c0 + t + i * NT;
        if (r >= p0 + LW && r < N) {
          const C* lrow = M + (long)r * ldm + p0;
#pragma unroll
          for (int b = 0; b < LW; ++b) {
            C l = lrow[b];
#pragma unroll
            for (int j = 0; j < LW; ++j) reg[i][j] = cfms(l, L.Xs[xb][b][j], reg[i][j]);
          }
        }
      }
      xb ^= 1;
    }
    // ---- factor the leaf (rows >= j0 are pivot candidates)
#pragma unroll
    for (int j = 0; j < LW; ++j) {
      if (j < w) {
        const int kc = j0 + j;
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
        const int tk = (kc - c0) % NT, sk = (kc - c0) / NT;
        if (t == tp) {
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (i == sp) {
#pragma unroll
              for (int jj = 0; jj < LW; ++jj) L.sP[jj] = reg[i][jj];
            }
        }
        if (p != kc && t == tk) {
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (i == sk) {
#pragma unroll
              for (int jj = 0; jj < LW; ++jj) L.sQ[jj] = reg[i][jj];
            }
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
          if (t == tk) {
#pragma unroll
            for (int i = 0; i < RPT; ++i)
              if (i == sk) {
#pragma unroll
                for (int jj = 0; jj < LW; ++jj) reg[i][jj] = L.sP[jj];
              }
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
          if (j < w) M[(long)r * ldm + j0 + j] = reg[i][j];
      }
    }
    // ---- unit-lower inverse of this leaf's diagonal block (for the later leaves of the panel)
    {
      const int a = t - (j0 - c0);
      if (a >= 0 && a < LW) {
#pragma unroll
        for (int j = 0; j < LW; ++j) L.Xs[xb][a][j] = reg[0][j];
      }
    }
    __syncthreads();
    if (t < LW) {  // column t of the inverse: y_t = 1, y_a = -sum_{t<=b<a} L_ab y_b
      for (int a = 0; a < LW; ++a) {
        C y = mk<T>(a == t ? 1 : 0, 0);
        if (a > t && a < w)
          for (int b = t; b < a; ++b) y = cfms(L.Xs[xb][a][b], L.Li[jl][b][t], y);
        L.Li[jl][a][t] = (a < w && t < w) ? y : mk<T>(0, 0);
      }
    }
    // ---- this leaf's interchanges on every other column of the rows
    bool any = false;
    for (int j = 0; j < w; ++j) any |= (L.spiv[j] != j0 + j);
    if (any) {
      for (int c = t; c < W; c += NT) {
        if (c >= j0 && c < j0 + w) continue;
        for (int j = 0; j < w; ++j) {
          int p = L.spiv[j];
          if (p != j0 + j) {
            long a0 = (long)(j0 + j) * ldm + c, a1 = (long)p * ldm + c;
            C u = M[a0];
            M[a0] = M[a1];
            M[a1] = u;
          }
        }
      }
    }
    xb ^= 1;
    __syncthreads();
  }
  // ---- inverses of the panel's L11 (unit lower) and U11 (upper) into the side buffer
  const int w = c1 - c0;
  for (int e = t; e < w * w; e += NT) {
    int a = e / w, b = e % w;
    L.u.v.Lb[a][b] = M[(long)(c0 + a) * ldm + c0 + b];
  }
  __syncthreads();
  if (t < w) {  // column t of L11^-1, from the top
    for (int a = 0; a < w; ++a) {
      C y = mk<T>(a == t ? 1 : 0, 0);
      if (a > t)
        for (int b = t; b < a; ++b) y = cfms(L.u.v.Lb[a][b], L.u.v.Xb[0][b][t], y);
      L.u.v.Xb[0][a][t] = (a >= t) ? y : mk<T>(0, 0);
    }
  } else if (t >= 64 && t < 64 + w) {  // column c of U11^-1, from the bottom
    const int c = t - 64;
    for (int a = w - 1; a >= 0; --a) {
      C y = mk<T>(0, 0);
      if (a <= c) {
        y = mk<T>(a == c ? 1 : 0, 0);
        for (int b = a + 1; b <= c; ++b) y = cfms(L.u.v.Lb[a][b], L.u.v.Xb[1][b][c], y);
        y = cdiv(y, L.u.v.Lb[a][a]);
      }
      L.u.v.Xb[1][a][c] = y;
    }
  }
  __syncthreads();
  C* dst = inv + (long)(c0 / P) * 2 * P * P;
  for (int e = t; e < 2 * P * P; e += NT) {
    int s = e / (P * P), a = (e / P) % P, b = e % P;
    dst[e] = (a < w && b < w) ? L.u.v.Xb[s][a][b] : mk<T>(0, 0);
  }
}

// ---------------------------------------------------------------- the kernel
template <typename T, int RPT, int LW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_lu(int N, int W, int ldm, int nops, const int* __restrict__ ops, cplx<T>* __restrict__ Mall,
     int* __restrict__ pivall, cplx<T>* __restrict__ invall, int* __restrict__ info) {
  __shared__ LuLds<T, LW> L;
  cplx<T>* M = Mall + (long)blockIdx.x * N * ldm;
  int* piv = pivall + (long)blockIdx.x * N;
  cplx<T>* inv = invall + (long)blockIdx.x * ((N + P - 1) / P) * 2 * P * P;
  int myinfo = 0;
  for (int o = 0; o < nops; ++o) {
    const int* op = ops + 8 * o;
    const int ty = op[0];
    if (ty == OP_PANEL) {
      wg_panel<T, RPT, LW>(M, ldm, N, W, op[1], op[2], piv, inv, L, myinfo);
    } else if (ty == OP_TRSM_L || ty == OP_TRSM_U) {
      const int k0 = op[1], k1 = op[2], c0 = op[3], c1 = op[4];
      const cplx<T>* iv = inv + (long)(k0 / P) * 2 * P * P + (ty == OP_TRSM_U ? P * P : 0);
      cplx<T>* X = M + (long)k0 * ldm + c0;
      wg_gemm<T, LW>(iv, P, X, ldm, X, ldm, k1 - k0, c1 - c0, k1 - k0, true, L);
    } else if (ty == OP_GEMM) {
      const int r0 = op[1], r1 = op[2], c0 = op[3], c1 = op[4], k0 = op[5], k1 = op[6];
      wg_gemm<T, LW>(M + (long)r0 * ldm + k0, ldm, M + (long)k0 * ldm + c0, ldm, M + (long)r0 * ldm + c0, ldm,
                     r1 - r0, c1 - c0, k1 - k0, false, L);
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

// splits keep every block boundary on a multiple of P
int half_split(int w) {
  int blocks = (w + P - 1) / P;
  return ((blocks + 1) / 2) * P;
}

void rec_trsm_l(Ops& ops, int k0, int k1, int c0, int c1) {
  if (c1 <= c0) return;
  if (k1 - k0 <= P) {
    ops.add(OP_TRSM_L, k0, k1, c0, c1);
    return;
  }
  int h = k0 + half_split(k1 - k0);
  rec_trsm_l(ops, k0, h, c0, c1);
  ops.add(OP_GEMM, h, k1, c0, c1, k0, h);
  rec_trsm_l(ops, h, k1, c0, c1);
}

void rec_trsm_u(Ops& ops, int k0, int k1, int c0, int c1) {
  if (c1 <= c0) return;
  if (k1 - k0 <= P) {
    ops.add(OP_TRSM_U, k0, k1, c0, c1);
    return;
  }
  int h = k0 + half_split(k1 - k0);
  rec_trsm_u(ops, h, k1, c0, c1);
  ops.add(OP_GEMM, k0, h, c0, c1, h, k1);
  rec_trsm_u(ops, k0, h, c0, c1);
}

// factor columns [c0, c1); `rhs` = the right-hand-side columns [N, W) ride along with this range
void rec_lu(Ops& ops, int N, int W, int c0, int c1, bool rhs) {
  if (c1 - c0 <= P) {
    ops.add(OP_PANEL, c0, c1, 0, 0);
    if (rhs && W > N) {
      rec_trsm_l(ops, c0, c1, N, W);
      if (c1 < N) ops.add(OP_GEMM, c1, N, N, W, c0, c1);
    }
    return;
  }
  int h = c0 + half_split(c1 - c0);
  rec_lu(ops, N, W, c0, h, false);
  int hi = rhs ? W : c1;
  rec_trsm_l(ops, c0, h, h, hi);
  ops.add(OP_GEMM, h, N, h, hi, c0, h);
  rec_lu(ops, N, W, h, c1, rhs);
}

}  // namespace

template <typename T, int RPT>
struct LeafW {
  static constexpr int value = 4;
};

static int round_rpt(int rpt) {
  static const int opts[] = {1, 2, 3, 4, 5, 6, 8, 12, 16};
  for (int o : opts)
    if (rpt <= o) return o;
  return -1;
}

// complex elements of the side buffer (panel inverses) per matrix
size_t lu_inv_elems(int N) { return (size_t)((N + P - 1) / P) * 2 * P * P; }

// factor + forward/backward solve of the appended right-hand sides
std::vector<int> lu_ops_factor(int N, int W) {
  Ops ops;
  rec_lu(ops, N, W, 0, N, true);
  if (W > N) rec_trsm_u(ops, 0, N, N, W);
  return ops.v;
}

// solve only, with the factors in place and new right-hand sides in columns [N, W)
std::vector<int> lu_ops_solve(int N, int W) {
  Ops ops;
  ops.add(OP_SWAP_RHS, 0, 0, 0, 0);
  rec_trsm_l(ops, 0, N, N, W);
  rec_trsm_u(ops, 0, N, N, W);
  return ops.v;
}

template <typename T>
bool launch_lu(int nmat, int N, int W, int ldm, const int* ops, int nops, cplx<T>* M, int* piv, cplx<T>* inv,
               int* info, hipStream_t s) {
  int rpt = round_rpt((N + NT - 1) / NT);
  dim3 grid(nmat), blk(NT);
#define ADGA_LU_CASE(R)                                                                                    \
  case R:                                                                                                  \
    hipLaunchKernelGGL((k_lu<T, R, LeafW<T, R>::value>), grid, blk, 0, s, N, W, ldm, nops, ops, M, piv, inv, \
                       info);                                                                              \
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
template bool launch_lu<float>(int, int, int, int, const int*, int, c32*, int*, c32*, int*, hipStream_t);
template bool launch_lu<double>(int, int, int, int, const int*, int, z64*, int*, z64*, int*, hipStream_t);

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
