// A7/A8: batched complex LU with partial pivoting, one workgroup per matrix, on gfx950 MFMA,
// and the streaming solve of the n^2 right-hand sides against the factors.
//
// M is row-major (row r, column c at r*ldm + c) with the matrix in columns [0, N); optional
// right-hand sides riding along in columns [N, W) are transformed with the rows (the explicit
// inverse uses that).  The factorisation is recursive (left half, Schur update of the right half,
// right half) down to panels of P columns, written out on the host as a list of operations so the
// kernel is a flat loop:
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
// Large triangular solves are recursive too, so the work runs as GEMMs.  The GEMM runs on MFMA
// (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64): a complex product is a real product with K
// doubled and re/im interleaved along it, the lane's K slot selecting (re, -im) or (im, re) of the
// A element and (re, im) of the B element.
#include <algorithm>
#include <vector>

#include "adga_kernels.h"
#include "mfma_gemm.h"

namespace adga {

constexpr int NT = 256;  // threads per workgroup (4 waves)
constexpr int P = 32;    // panel width: the recursion's leaf, and the TRSM block

enum { OP_PANEL = 0, OP_TRSM_L = 1, OP_TRSM_U = 2, OP_GEMM = 3 };

// One GEMM or TRSM operation on every matrix of the batch, one tile per workgroup.
//   mode 0: M[r0:r0+m, c0:c0+n] -= M[r0:r0+m, k0:k0+K] M[k0:k0+K, c0:c0+n]
//   mode 1/2: M[k0:k0+K, c0:c0+n] <- D M[k0:k0+K, c0:c0+n] with D the stored inverse of the unit-lower
//             (1) or upper (2) diagonal block of panel k0 (m = K <= P, in place)
// Workgroups are dealt round-robin to the 8 XCDs; the decode keeps all tiles of one matrix on one
// XCD and adjacent in time, so its A row panels and B column panels are shared through that L2.
template <typename T, int TMT, int TNT, int WR, int WC>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) k_gemm(int nmat, int ntc, int ntile,
                                                                                  long mstride, cplx<T>* __restrict__ Mall,
                                             int ldm, int r0, int m, int c0, int n, int k0, int K, int mode,
                                             const cplx<T>* __restrict__ invall, long istride) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[RmCfg<T, TMT, TNT>::LDS];
  const int L = blockIdx.x, x = L & 7, sl = L >> 3;
  const int mat = (sl / ntile) * 8 + x, tile = sl % ntile;
  if (mat >= nmat) return;
  const int rt = (tile / ntc) * TMT, ct = (tile % ntc) * TNT;
  cplx<T>* M = Mall + (long)mat * mstride;
  if (mode == 0) {
    gemm_tile<T, TMT, TNT, WR, WC>(M + (long)r0 * ldm + k0, ldm, M + (long)k0 * ldm + c0, ldm,
                                   M + (long)r0 * ldm + c0, ldm, m, n, K, false, rt, ct, lds);
  } else {
    const cplx<T>* D = invall + (long)mat * istride + (long)(k0 / P) * 2 * P * P + (mode == 2 ? P * P : 0);
    cplx<T>* X = M + (long)k0 * ldm + c0;
    gemm_tile<T, TMT, TNT, WR, WC>(D, P, X, ldm, X, ldm, K, n, K, true, rt, ct, lds);
  }
}

// The Schur updates (mode 0) with 64- and 128-wide tiles, by the 3M product: eight waves, two
// K-chunk buffers in LDS, the same XCD-aware tile decode as k_gemm.
template <typename T, int TMT, int TNT, int WR, int WC, int KC>
__global__ void __launch_bounds__(WR * WC * 64) k_gemm3m(int nmat, int ntc, int ntile, long mstride,
                                                         cplx<T>* __restrict__ Mall, int ldm, int r0, int m, int c0,
                                                         int n, int k0, int K) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[G3Cfg<T, TMT, TNT, WR, WC, KC>::LDS];
  const int L = blockIdx.x, x = L & 7, sl = L >> 3;
  const int mat = (sl / ntile) * 8 + x, tile = sl % ntile;
  if (mat >= nmat) return;
  const int rt = (tile / ntc) * TMT, ct = (tile % ntc) * TNT;
  cplx<T>* M = Mall + (long)mat * mstride;
  cplx<T>* Cp = M + (long)r0 * ldm + c0;
  gemm3m_tile<T, TMT, TNT, WR, WC, KC>(
      M + (long)r0 * ldm + k0, ldm, M + (long)k0 * ldm + c0, ldm, [=](int r, int c) { return Cp[r * ldm + c]; },
      [=](int r, int c, cplx<T> v) { Cp[r * ldm + c] = v; }, m, n, K, false, true, rt, ct, lds);
}

// component-wise select of two complex values (a select of whole structs can turn into a select of
// their addresses, which forces register arrays into scratch memory)
template <typename T>
__device__ __forceinline__ cplx<T> csel(bool c, cplx<T> a, cplx<T> b) {
  return mk<T>(c ? a.re : b.re, c ? a.im : b.im);
}

// Lane moves inside 16-lane rows (DPP row_ror) and reductions over a row: every lane of the row
// ends with the row's result.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_v(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_v(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i<CTRL>((int)b), hi = dpp_i<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename T>
__device__ __forceinline__ T row_max(T v) {  // no NaN among the values
  T o = dpp_v<0x121>(v);
  v = o > v ? o : v;
  o = dpp_v<0x122>(v);
  v = o > v ? o : v;
  o = dpp_v<0x124>(v);
  v = o > v ? o : v;
  o = dpp_v<0x128>(v);
  return o > v ? o : v;
}
__device__ __forceinline__ int row_min(int v) {
  v = min(v, dpp_i<0x121>(v));
  v = min(v, dpp_i<0x122>(v));
  v = min(v, dpp_i<0x124>(v));
  return min(v, dpp_i<0x128>(v));
}
template <typename T>
__device__ __forceinline__ T lane_value(T v, int l);
template <>
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <>
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// the wave's maximum, the same value in every lane
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  v = row_max(v);
  const T a = lane_value(v, 0), b = lane_value(v, 16), c = lane_value(v, 32), d = lane_value(v, 48);
  const T ab = a > b ? a : b, cd = c > d ? c : d;
  return ab > cd ? ab : cd;
}

// Panel LDS (NW waves)
template <typename T, int LW, int NW>
struct PanelLds {
  struct {
    struct {
      cplx<T> Lb[P][P + 1];     // the panel's diagonal block
      cplx<T> Xb[2][P][P + 1];  // its inverses being formed (L, U)
    } v;
  } u;
  // pivot search, double-buffered by column parity: each wave's best candidate (|re|+|im|, row) with
  // the row's leaf values, and the values of the row the pivot replaces
  cplx<T> cv[2][NW][LW];
  T cval[2][NW];
  int crow[2][NW];
  cplx<T> kv[2][LW];
  cplx<T> Xs[2][LW][LW + 1];       // left-looking: [0] raw rows of the current leaf, [1] their U block
  cplx<T> Li[P / LW][LW][LW + 1];  // unit-lower inverses of the panel's leaves
  int spiv[LW];
};

// LW consecutive complex of one row with 16-byte accesses (the row slice is 16-byte aligned: the
// matrix rows and the leaf columns start on multiples of 16 bytes)
template <typename T, int LW>
struct RowSlice {
  static constexpr int PER = 16 / (int)sizeof(cplx<T>);
  static_assert(LW % PER == 0, "leaf width in 16-byte units");
  struct alignas(16) V { cplx<T> v[PER]; };
  static __device__ __forceinline__ void load(const cplx<T>* p, cplx<T> (&d)[LW]) {
#pragma unroll
    for (int u = 0; u < LW / PER; ++u) {
      const V x = reinterpret_cast<const V*>(p)[u];
#pragma unroll
      for (int q = 0; q < PER; ++q) d[u * PER + q] = x.v[q];
    }
  }
  static __device__ __forceinline__ void store(cplx<T>* p, const cplx<T> (&d)[LW]) {
#pragma unroll
    for (int u = 0; u < LW / PER; ++u) {
      V x;
#pragma unroll
      for (int q = 0; q < PER; ++q) x.v[q] = d[u * PER + q];
      reinterpret_cast<V*>(p)[u] = x;
    }
  }
};

// ---------------------------------------------------------------- panel factorisation
// Columns [c0, c1) of rows [c0, N) with NTP threads; thread t holds rows c0 + t + i*NTP (i < RPT)
// of the current leaf of LW columns (rows >= N as zeros).  Barriers: two per (leaf, earlier leaf)
// pair of the left-looking updates, one per column for the pivot (each wave publishes its best
// candidate together with that row's values, so the winner's row needs no second round), one per
// leaf.  Row updates are written without branches on the row (a multiplier of zero leaves a row
// as it is); a slot whose rows are all outside the update for the whole wave is skipped on a
// wave-uniform test.
template <typename T, int RPT, int LW, int NTP>
__device__ __forceinline__ void wg_panel(cplx<T>* __restrict__ M, int ldm, int N, int W, int c0, int c1,
                                         int* __restrict__ piv, cplx<T>* __restrict__ inv,
                                         PanelLds<T, LW, NTP / 64>& L, int& myinfo) {
  using C = cplx<T>;
  constexpr int NW = NTP / 64;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int rb = c0 + t;            // this thread's row in slot 0
  const int wr = c0 + wave * 64;    // the wave's first row in slot 0 (uniform)
  const int nl = (c1 - c0 + LW - 1) / LW;
  int q = 0;  // parity of the pivot-search buffers
  for (int jl = 0; jl < nl; ++jl) {
    const int j0 = c0 + jl * LW;
    const int w = min(LW, c1 - j0);
    C reg[RPT][LW];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = rb + i * NTP;
      if (r < N && w == LW) {
        RowSlice<T, LW>::load(M + (long)r * ldm + j0, reg[i]);
      } else {
#pragma unroll
        for (int j = 0; j < LW; ++j) {
          C v = mk<T>(0, 0);
          if (r < N && j < w) v = M[(long)r * ldm + j0 + j];
          reg[i][j] = v;
        }
      }
    }
    // ---- left-looking updates from the panel's previous leaves
    for (int il = 0; il < jl; ++il) {
      const int p0 = c0 + il * LW;
      const int a = t - (p0 - c0);  // this thread owns row p0 + a (slot 0) when 0 <= a < LW
      const bool owner = a >= 0 && a < LW;
      C lrow[RPT][LW];  // L values of the rows below the earlier leaf (zero elsewhere), loaded ahead
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = rb + i * NTP;
        if (r >= p0 + LW && r < N) {
          RowSlice<T, LW>::load(M + (long)r * ldm + p0, lrow[i]);
        } else {
#pragma unroll
          for (int b = 0; b < LW; ++b) lrow[i][b] = mk<T>(0, 0);
        }
      }
      if (owner) {
#pragma unroll
        for (int j = 0; j < LW; ++j) L.Xs[0][a][j] = reg[0][j];
      }
      __syncthreads();
      if (t < LW * LW) {  // U block element (a2, jc): L_ii^-1 applied to the raw rows
        const int a2 = t / LW, jc = t % LW;
        C acc = mk<T>(0, 0);
#pragma unroll
        for (int b = 0; b < LW; ++b) acc = cfma(L.Li[il][a2][b], L.Xs[0][b][jc], acc);
        L.Xs[1][a2][jc] = acc;
      }
      __syncthreads();
      if (owner) {
#pragma unroll
        for (int j = 0; j < LW; ++j) reg[0][j] = L.Xs[1][a][j];
      }
#pragma unroll
      for (int b = 0; b < LW; ++b) {
        C u[LW];
#pragma unroll
        for (int j = 0; j < LW; ++j) u[j] = L.Xs[1][b][j];
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          if (wr + i * NTP + 63 < p0 + LW || wr + i * NTP >= N) continue;
#pragma unroll
          for (int j = 0; j < LW; ++j) reg[i][j] = cfms(lrow[i][b], u[j], reg[i][j]);
        }
      }
    }
    // ---- factor the leaf (rows >= kc are pivot candidates).  Step j resolves column j's pivot
    // and multipliers, updates column j + 1, runs column j + 1's pivot search, and only then
    // updates the columns beyond j + 1 (that update fills the latency of the search's reductions);
    // the candidate rows are published whole after it.  Step -1 only searches column 0.
#pragma unroll
    for (int j = -1; j < LW; ++j) {
      if (j >= w) continue;
      C pr[LW], lm[RPT];
      if (j >= 0) {
        const int kc = j0 + j;
        __syncthreads();
        // the workgroup's pivot: the waves' candidates in lanes 0 .. NW-1 (one DPP row); ties in
        // value go to the smallest row
        T gv = (T)-1;
        int grow = 0x7fffffff;
        if (lane < NW) {
          gv = L.cval[q][lane];
          grow = L.crow[q][lane];
        }
        const T gm = lane_value(row_max(gv), 0);
        const unsigned long long tied = __ballot(lane < NW && gm >= (T)0 && gv == gm);
        int gr, pw;
        if (__popcll(tied) <= 1) {
          pw = tied ? __ffsll((long long)tied) - 1 : 0;
          gr = __builtin_amdgcn_readlane(grow, pw);
        } else {
          gr = __builtin_amdgcn_readlane(row_min((gm >= (T)0 && gv == gm) ? grow : 0x7fffffff), 0);
          const unsigned long long pb = __ballot(lane < NW && gm >= (T)0 && gv == gm && grow == gr);
          pw = __ffsll((long long)pb) - 1;
        }
        int p = gm >= (T)0 ? gr : N;
        // no finite candidate (NaN column): keep the diagonal row and flag the matrix
        if (p >= N) {
          p = kc;
          if (myinfo == 0) myinfo = kc + 1;
        }
        const C* src = (p == kc) ? &L.kv[q][0] : &L.cv[q][pw][0];
#pragma unroll
        for (int jj = 0; jj < LW; ++jj) pr[jj] = src[jj];
        if (t == 0) { L.spiv[j] = p; piv[kc] = p; }
        if (p != kc) {  // the interchange, in registers: row p gets the old row kc and vice versa
          // (selected by row, not by slot index: an index into reg would move it to scratch memory)
          C kvv[LW];
#pragma unroll
          for (int jj = 0; jj < LW; ++jj) kvv[jj] = L.kv[q][jj];
#pragma unroll
          for (int i = 0; i < RPT; ++i) {
            const int r = rb + i * NTP;
            const bool hp = r == p, hk = r == kc;
#pragma unroll
            for (int jj = 0; jj < LW; ++jj) reg[i][jj] = csel(hp, kvv[jj], csel(hk, pr[jj], reg[i][jj]));
          }
        }
        const C pv = pr[j];
        const bool nz = !(pv.re == (T)0 && pv.im == (T)0);
        if (!nz && myinfo == 0) myinfo = kc + 1;  // zero pivot: no elimination in this column
        const C rp = crecip(pv);
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const bool below = nz && rb + i * NTP > kc;  // rows >= N hold zeros and stay zero
          const C l = cmul(reg[i][j], rp);
          lm[i] = csel(below, l, mk<T>(0, 0));
          reg[i][j] = csel(below, l, reg[i][j]);
        }
        q ^= 1;
      }
      const int jn = j + 1;
      int wrow = -1, wslot = 0, wlane = 0;  // this wave's candidate for column jn
      if (jn < LW && jn < w) {
        const int kn = j0 + jn;
        if (j >= 0) {
#pragma unroll
          for (int i = 0; i < RPT; ++i) reg[i][jn] = cfms(lm[i], pr[jn], reg[i][jn]);
        }
        // this thread's best candidate for column jn (|re|+|im|, -1 for none; the first slot on ties)
        T bvv = (T)-1;
        int bs = 0;
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const int r = rb + i * NTP;
          const T v = cabs1(reg[i][jn]);
          const T vc = (r >= kn && r < N && v == v) ? v : (T)-1;
          const bool take = vc > bvv;
          bvv = take ? vc : bvv;
          bs = take ? i : bs;
        }
        // the wave's best value, and its smallest row (slot first, then lane): uniform
        const T wm = wave_max(bvv);
        if (wm >= (T)0) {
#pragma unroll
          for (int i = 0; i < RPT; ++i) {
            if (wrow < 0) {
              const unsigned long long b = __ballot(bvv == wm && bs == i);
              if (b != 0ull) {
                wlane = __ffsll((long long)b) - 1;
                wslot = i;
                wrow = wr + wlane + i * NTP;
              }
            }
          }
        }
        if (lane == 0) {
          L.cval[q][wave] = wm;
          L.crow[q][wave] = wrow;
        }
      }
      if (j >= 0) {
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
#pragma unroll
          for (int jj = j + 2; jj < LW; ++jj) reg[i][jj] = cfms(lm[i], pr[jj], reg[i][jj]);
        }
      }
      if (jn < LW && jn < w) {  // rows published whole, after the update of all their columns
        if (wrow >= 0 && lane == wlane) {  // the row's leaf values go with the candidate
#pragma unroll
          for (int i = 0; i < RPT; ++i)
            if (i == wslot) {
#pragma unroll
              for (int jj = 0; jj < LW; ++jj) L.cv[q][wave][jj] = reg[i][jj];
            }
        }
        if (t == j0 + jn - c0) {  // the diagonal row is slot 0 of thread kn - c0 (kn - c0 < P)
#pragma unroll
          for (int jj = 0; jj < LW; ++jj) L.kv[q][jj] = reg[0][jj];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = rb + i * NTP;
      if (r < N) {
        if (w == LW) {
          RowSlice<T, LW>::store(M + (long)r * ldm + j0, reg[i]);
        } else {
#pragma unroll
          for (int j = 0; j < LW; ++j)
            if (j < w) M[(long)r * ldm + j0 + j] = reg[i][j];
        }
      }
    }
    // ---- unit-lower inverse of this leaf's diagonal block (for the later leaves of the panel)
    {
      const int a = t - (j0 - c0);
      if (a >= 0 && a < LW) {
#pragma unroll
        for (int j = 0; j < LW; ++j) L.Xs[0][a][j] = reg[0][j];
      }
    }
    __syncthreads();
    if (t < LW) {  // column t of the inverse: y_a = [a == t] - sum_{b < a} L_ab y_b (zero above t)
      C y[LW];
#pragma unroll
      for (int a = 0; a < LW; ++a) {
        C acc = mk<T>(a == t ? 1 : 0, 0);
#pragma unroll
        for (int b = 0; b < a; ++b) acc = cfms(L.Xs[0][a][b], y[b], acc);
        y[a] = csel(a < w && t < w, acc, mk<T>(0, 0));
        L.Li[jl][a][t] = y[a];
      }
    }
    // ---- this leaf's interchanges on every other column of the rows
    bool any = false;
    for (int j = 0; j < w; ++j) any |= (L.spiv[j] != j0 + j);
    if (any) {
      for (int c = t; c < W; c += NTP) {
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
    __syncthreads();
  }
  // ---- L11^-1 (unit lower) into the side buffer, by block recurrence on the leaves' inverses:
  // X_ii = Li[i], X_ij = -Li[i] sum_{j <= k < i} L_ik X_kj, one block diagonal at a time (U11^-1 is
  // formed after the factorisation by k_lu_uinv, off this critical path)
  const int w = c1 - c0;
  const int nlf = (w + LW - 1) / LW;
  for (int e = t; e < P * P; e += NTP) {
    const int a = e / P, b = e % P, bi = a / LW, bj = b / LW;
    L.u.v.Lb[a][b] = (a < w && b < w) ? M[(long)(c0 + a) * ldm + c0 + b] : mk<T>(0, 0);
    L.u.v.Xb[0][a][b] = (bi == bj && bi < nlf) ? L.Li[bi][a % LW][b % LW] : mk<T>(0, 0);
  }
  __syncthreads();
  for (int dd = 1; dd < nlf; ++dd) {
    const int nb = nlf - dd;
    for (int e = t; e < nb * LW * LW; e += NTP) {
      const int j = e / (LW * LW), i = j + dd, a = (e / LW) % LW, bb = e % LW;
      C acc = mk<T>(0, 0);
      for (int k = j * LW; k < i * LW; ++k) acc = cfma(L.u.v.Lb[i * LW + a][k], L.u.v.Xb[0][k][j * LW + bb], acc);
      L.u.v.Xb[1][i * LW + a][j * LW + bb] = acc;
    }
    __syncthreads();
    for (int e = t; e < nb * LW * LW; e += NTP) {
      const int j = e / (LW * LW), i = j + dd, a = (e / LW) % LW, bb = e % LW;
      C acc = mk<T>(0, 0);
#pragma unroll
      for (int c = 0; c < LW; ++c) acc = cfma(L.Li[i][a][c], L.u.v.Xb[1][i * LW + c][j * LW + bb], acc);
      L.u.v.Xb[0][i * LW + a][j * LW + bb] = mk<T>(-acc.re, -acc.im);
    }
    __syncthreads();
  }
  C* dst = inv + (long)(c0 / P) * 2 * P * P;
  for (int e = t; e < P * P; e += NTP) {
    const int a = e / P, b = e % P;
    dst[e] = (a < w && b < w) ? L.u.v.Xb[0][a][b] : mk<T>(0, 0);
  }
}

// U11^-1 of every panel of every matrix (grid: panels x matrices), after the factorisation: column c
// of the inverse by back substitution in lane c, fully unrolled so the column stays in registers
// and U11 is read from LDS by all lanes at once; y_a = 0 for a > c follows from the recurrence
// itself.  Padding rows of a narrow last panel get a unit diagonal.
template <typename T>
__global__ void __launch_bounds__(64) k_lu_uinv(int N, int ldm, long mstride, long istride,
                                                const cplx<T>* __restrict__ Mall, cplx<T>* __restrict__ invall) {
  using C = cplx<T>;
  __shared__ C Ub[P][P + 1];
  __shared__ C rd[P];
  const int t = threadIdx.x, c0 = blockIdx.x * P, w = min(P, N - c0);
  const C* M = Mall + (long)blockIdx.y * mstride;
  for (int e = t; e < P * P; e += 64) {
    const int a = e / P, b = e % P;
    Ub[a][b] = (a < w && b < w) ? M[(long)(c0 + a) * ldm + c0 + b] : mk<T>(a == b ? 1 : 0, 0);
  }
  __syncthreads();
  if (t < P) rd[t] = crecip(Ub[t][t]);
  __syncthreads();
  if (t < P) {
    const int c = t;
    C y[P];
#pragma unroll
    for (int a = P - 1; a >= 0; --a) {
      C acc = mk<T>(a == c ? 1 : 0, 0);
#pragma unroll
      for (int b = a + 1; b < P; ++b) acc = cfms(Ub[a][b], y[b], acc);
      y[a] = cmul(acc, rd[a]);
    }
    C* dst = invall + (long)blockIdx.y * istride + (long)blockIdx.x * 2 * P * P + P * P;
#pragma unroll
    for (int a = 0; a < P; ++a) dst[a * P + c] = (a < w && c < w) ? y[a] : mk<T>(0, 0);
  }
}

// ---------------------------------------------------------------- panel, pivots
// PANEL(c0, c1) of every matrix of the batch, one workgroup of NTP threads per matrix.  info[m]
// keeps the first zero pivot (1-based, LAPACK's info), or stays 0.
template <typename T, int RPT, int LW, int NTP>
__global__ void __launch_bounds__(NTP) k_panel(int N, int W, int ldm, long mstride, int c0, int c1,
                                               cplx<T>* __restrict__ Mall, int* __restrict__ pivall,
                                               cplx<T>* __restrict__ invall, long istride, int* __restrict__ info) {
  __shared__ PanelLds<T, LW, NTP / 64> L;
  const int mat = blockIdx.x;
  int myinfo = 0;
  wg_panel<T, RPT, LW, NTP>(Mall + (long)mat * mstride, ldm, N, W, c0, c1, pivall + (long)mat * N,
                            invall + (long)mat * istride, L, myinfo);
  if (threadIdx.x == 0 && myinfo != 0 && info[mat] == 0) info[mat] = myinfo;
}

__global__ void k_lu_init(int nmat, int* __restrict__ info) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nmat) info[i] = 0;
}

// The pivot sequence in `piv` (LAPACK's ipiv, 0-based) is replaced by the row permutation it
// amounts to: row j of the factors is row piv[j] of the input, which is what the solve gathers with.
constexpr int PERM_MAX_N = 4096;
__global__ void __launch_bounds__(NT) k_lu_perm(int N, int* __restrict__ pivall) {
  __shared__ int pl[PERM_MAX_N], ps[PERM_MAX_N];
  int* piv = pivall + (long)blockIdx.x * N;
  for (int j = threadIdx.x; j < N; j += NT) {
    pl[j] = j;
    ps[j] = piv[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < N; ++j) {
      const int p = ps[j];
      if (p != j) {
        const int u = pl[j];
        pl[j] = pl[p];
        pl[p] = u;
      }
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < N; j += NT) piv[j] = pl[j];
}

// ---------------------------------------------------------------- the solve kernel
// Y (+)= (P^T L U)^-1 R for the ns right-hand-side columns [g*ns, ...) of matrix m, R and Y in FP64 with
// layout [matrix][row][nrhs].  When the columns do not fit the LDS at once they are split into ng groups,
// each a workgroup that streams all of the matrix's factors; the groups of one matrix are therefore
// dispatched on one XCD in consecutive slots (workgroup b runs on XCD b % 8), so that they read the same
// factor rows at about the same time through that XCD's L2 instead of each from HBM.  The columns are gathered through
// the permutation into LDS (in the factor precision) and stay there; the factors stream from HBM
// once: block row by block row (P rows), the off-diagonal part as dot products against the
// resident solution (16 lanes per row, shuffle reduction), the diagonal block through its stored
// inverse.  With 512 threads wave w owns rows 4w .. 4w+3 of a block row, lane l row 4w + l/16 at columns
// l%16 + 16j (with 256 threads: rows 8w .. 8w+7, lane l rows 8w + l/16 and 8w + 4 + l/16).
constexpr int SOLVE_LDS_BYTES = 80896;        // two workgroups per CU
constexpr int SOLVE_LDS_BYTES_WIDE = 163840;  // one workgroup per CU (the whole LDS)

template <typename T, int CG, int LDSB, int NTS>
__global__ void __launch_bounds__(NTS) k_lu_solve(int N, int nrhs, int ns, int ng, int nmat, int xcd, int ldm,
                                                const cplx<T>* __restrict__ Mall, const int* __restrict__ permall,
                                                const cplx<T>* __restrict__ invall, const z64* __restrict__ R,
                                                z64* __restrict__ Y, int accumulate, const int* __restrict__ active,
                                                int from_m) {
  using C = cplx<T>;
  // x[j * CG + s]: rows [N, N + P) and columns [nsg, CG) are zeros, so every loop below runs over CG
  // columns without predicates
  __shared__ C x[LDSB / sizeof(C)];
  // ROWS rows per 16-lane group: 4 waves x 8 rows (256 threads) or 8 waves x 4 rows (512) cover a block row
  constexpr int ROWS = NTS == 256 ? 2 : 1;
  static_assert(NTS == 256 || NTS == 512, "256 or 512 threads");
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, rs = lane >> 4, ks = lane & 15;
  // xcd: the XCD-grouped mapping; else workgroup b is group b % ng of matrix b / ng (few matrices: their
  // groups spread over every XCD)
  const int slot = xcd ? blockIdx.x >> 3 : blockIdx.x;
  const int m = xcd ? (slot / ng) * 8 + (blockIdx.x & 7) : slot / ng, grp = slot % ng;
  if (m >= nmat) return;  // the grid is rounded up to whole XCD rounds: the whole workgroup leaves
  const int s0 = grp * ns, nsg = min(ns, nrhs - s0);
  if (active && !active[m]) return;  // the whole workgroup leaves: no barrier is reached
  const C* M = Mall + (long)m * N * ldm;
  const int* perm = permall + (long)m * N;
  const int nb = (N + P - 1) / P;
  const C* inv = invall + (long)m * nb * 2 * P * P;
  const z64* Rm = R ? R + (long)m * N * nrhs + s0 : nullptr;  // not read with from_m
  z64* Ym = Y + (long)m * N * nrhs + s0;
  if (from_m) {
    // the right-hand sides the factorisation carried along: L^-1 P B in the columns [N, N + nrhs) of M
    for (int e = t; e < (N + P) * CG; e += NTS) {
      const int j = e / CG, s = e % CG;
      x[e] = (j < N && s < nsg) ? M[(long)j * ldm + N + s0 + s] : mk<T>(0, 0);
    }
  } else {
    for (int e = t; e < (N + P) * CG; e += NTS) {
      const int j = e / CG, s = e % CG;
      x[e] = (j < N && s < nsg) ? ccast<T>(Rm[(long)perm[j] * nrhs + s]) : mk<T>(0, 0);
    }
  }
  __syncthreads();
  constexpr int SU = 8, G = 16 * SU;  // loads per row per group; k covered by one group of 16 lanes
  for (int pass = from_m ? 1 : 0; pass < 2; ++pass) {  // 0: forward with unit-lower L, 1: backward with U
    for (int q = 0; q < nb; ++q) {
      const int ib = pass == 0 ? q : nb - 1 - q;
      const int r0 = ib * P, r1 = min(N, r0 + P);
      const int k0 = pass == 0 ? 0 : r1, k1 = pass == 0 ? r0 : N;
      if (k1 > k0) {
        const int ra = r0 + wave * 4 * ROWS + rs, rb = ROWS == 2 ? ra + 4 : N;  // rb unused with one row
        const C* La = M + (long)min(ra, N - 1) * ldm;
        const C* Lb = M + (long)min(rb, N - 1) * ldm;
        C acc0[CG], acc1[CG];
#pragma unroll
        for (int s = 0; s < CG; ++s) acc0[s] = acc1[s] = mk<T>(0, 0);
        C pa0[SU], pa1[SU], qa0[SU], qa1[SU];
        auto load = [&](C* d0, C* d1, int kb) {
#pragma unroll
          for (int u = 0; u < SU; ++u) {
            d0[u] = La[kb + 16 * u];
            if constexpr (ROWS == 2) d1[u] = Lb[kb + 16 * u];
          }
        };
        auto use = [&](const C* d0, const C* d1, int kb) {
#pragma unroll
          for (int u = 0; u < SU; ++u) {
            const C* xr = x + (kb + 16 * u) * CG;
#pragma unroll
            for (int s = 0; s < CG; ++s) {
              const C xv = xr[s];
              acc0[s] = cfma(d0[u], xv, acc0[s]);
              if constexpr (ROWS == 2) acc1[s] = cfma(d1[u], xv, acc1[s]);
            }
          }
        };
        // full groups (every lane's SU columns inside [k0, k1)): software-pipelined, no predicates
        const int nfull = (k1 - k0) / G;
        int k = k0 + ks;
        if (nfull > 0) {
          load(pa0, pa1, k);
          for (int g = 0; g < nfull; g += 2) {
            if (g + 1 < nfull) load(qa0, qa1, k + G);
            use(pa0, pa1, k);
            if (g + 1 >= nfull) {
              k += G;
              break;
            }
            if (g + 2 < nfull) load(pa0, pa1, k + 2 * G);
            use(qa0, qa1, k + G);
            k += 2 * G;
          }
        }
        if (k < k1) {  // the rest, fewer than G columns: one predicated group, its loads in flight together
#pragma unroll
          for (int u = 0; u < SU; ++u) {
            const bool in = k + 16 * u < k1;
            pa0[u] = in ? La[k + 16 * u] : mk<T>(0, 0);
            if constexpr (ROWS == 2) pa1[u] = in ? Lb[k + 16 * u] : mk<T>(0, 0);
          }
#pragma unroll
          for (int u = 0; u < SU; ++u) {
            if (k + 16 * u < k1) {
              const C* xr = x + (k + 16 * u) * CG;
#pragma unroll
              for (int s = 0; s < CG; ++s) {
                const C xv = xr[s];
                acc0[s] = cfma(pa0[u], xv, acc0[s]);
                if constexpr (ROWS == 2) acc1[s] = cfma(pa1[u], xv, acc1[s]);
              }
            }
          }
        }
#pragma unroll
        for (int s = 0; s < CG; ++s) {
#pragma unroll
          for (int off = 8; off > 0; off >>= 1) {
            acc0[s].re += __shfl_xor(acc0[s].re, off);
            acc0[s].im += __shfl_xor(acc0[s].im, off);
            if constexpr (ROWS == 2) {
              acc1[s].re += __shfl_xor(acc1[s].re, off);
              acc1[s].im += __shfl_xor(acc1[s].im, off);
            }
          }
        }
        if (ks == 0) {
#pragma unroll
          for (int s = 0; s < CG; ++s) {
            if (ra < r1) x[ra * CG + s] = csub(x[ra * CG + s], acc0[s]);
            if (ROWS == 2 && rb < r1) x[rb * CG + s] = csub(x[rb * CG + s], acc1[s]);
          }
        }
      }
      __syncthreads();
      // diagonal block: x_b <- D^-1 x_b with D^-1 the stored inverse (P x P, zero beyond the block)
      const C* Di = inv + (long)ib * 2 * P * P + (pass == 0 ? 0 : P * P);
      constexpr int H = (P * CG + NTS - 1) / NTS;  // entries of the block per thread
      C o[H];
#pragma unroll
      for (int h = 0; h < H; ++h) {
        const int e = t + h * NTS;
        o[h] = mk<T>(0, 0);
        if (e < P * CG) {
          const int r = e / CG, s = e % CG;
          if (r0 + r < r1) {
            // the whole row of D^-1 (zero-padded to P) in one round of loads
            const C* drow = Di + r * P;
            const int c0 = pass == 0 ? 0 : r, c1 = pass == 0 ? r + 1 : r1 - r0;
            C dv[P];
#pragma unroll
            for (int c = 0; c < P; ++c) dv[c] = drow[c];
            C acc = mk<T>(0, 0);
#pragma unroll
            for (int c = 0; c < P; ++c) {
              const bool ok = c >= c0 && c < c1;
              acc = cfma(ok ? dv[c] : mk<T>(0, 0), x[(r0 + c) * CG + s], acc);
            }
            o[h] = acc;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int h = 0; h < H; ++h) {
        const int e = t + h * NTS;
        if (e < P * CG) {
          const int r = e / CG;
          if (r0 + r < r1) x[r0 * CG + e] = o[h];
        }
      }
      __syncthreads();
    }
  }
  for (int e = t; e < N * nsg; e += NTS) {
    const int j = e / nsg, s = e % nsg;
    const C v = x[j * CG + s];
    z64 y = mk((double)v.re, (double)v.im);
    if (accumulate) y = cadd(Ym[(long)j * nrhs + s], y);
    Ym[(long)j * nrhs + s] = y;
  }
}

// ---------------------------------------------------------------- host: operation lists
namespace {

// operation list: a growing vector, or (buf non-null) a caller's fixed array of cap ints, in which
// case n still counts every int the list needs
struct Ops {
  std::vector<int> v;
  int* buf = nullptr;
  size_t cap = 0, n = 0;
  void add(int ty, int a, int b, int c, int d, int e = 0, int f = 0) {
    int o[8] = {ty, a, b, c, d, e, f, 0};
    if (!buf) v.insert(v.end(), o, o + 8);
    else if (n + 8 <= cap)
      for (int i = 0; i < 8; ++i) buf[n + i] = o[i];
    n += 8;
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

// leaf width: 8 columns in registers for complex64 up to 6 rows per thread, else 4
template <typename T, int RPT>
struct LeafW {
  static constexpr int value = (sizeof(T) == 4 && RPT <= 6) ? 8 : 4;
};

// panel shape by the panel's row count (rows c0 .. N-1): 256 threads up to 1536 rows, else 512,
// with the fewest rows per thread (RPT) that cover them, encoded as RPT + 100 for the 512-thread
// form; -1 beyond 4096 rows.  Fitting RPT to each panel keeps the per-column work of the later,
// shorter panels proportional to their rows.
static int panel_cfg(int rows) {
  if (rows <= 1536) return std::max(1, (rows + 255) / 256);
  static const int opts[] = {4, 5, 6, 8};
  const int rpt = (rows + 511) / 512;
  for (int o : opts)
    if (rpt <= o) return 100 + o;
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

// factor + forward elimination of the appended right-hand sides (L^-1 P B in columns [N, W); the backward
// substitution is left to the solve kernel, which then reads only U)
std::vector<int> lu_ops_factor_fwd(int N, int W) {
  Ops ops;
  rec_lu(ops, N, W, 0, N, true);
  return ops.v;
}

size_t lu_ops_factor_into(int N, int W, int* buf, size_t cap) {
  Ops ops;
  ops.buf = buf;
  ops.cap = cap;
  rec_lu(ops, N, W, 0, N, true);
  if (W > N) rec_trsm_u(ops, 0, N, N, W);
  return ops.n;
}

namespace {

template <typename T, int TMT, int TNT, int WR, int WC>
void gemm_launch(int nmat, long mstride, cplx<T>* M, int ldm, int r0, int m, int c0, int n, int k0, int K, int mode,
                 const cplx<T>* inv, long istride, hipStream_t s) {
  const int ntr = (m + TMT - 1) / TMT, ntc = (n + TNT - 1) / TNT, ntile = ntr * ntc;
  const long nwg = (long)ntile * ((nmat + 7) / 8) * 8;
  hipLaunchKernelGGL((k_gemm<T, TMT, TNT, WR, WC>), dim3((unsigned)nwg), dim3(NT), 0, s, nmat, ntc, ntile, mstride, M,
                     ldm, r0, m, c0, n, k0, K, mode, inv, istride);
}

template <typename T, int TMT, int TNT, int WR, int WC, int KC>
void gemm3m_launch(int nmat, long mstride, cplx<T>* M, int ldm, int r0, int m, int c0, int n, int k0, int K,
                   hipStream_t s) {
  const int ntr = (m + TMT - 1) / TMT, ntc = (n + TNT - 1) / TNT, ntile = ntr * ntc;
  const long nwg = (long)ntile * ((nmat + 7) / 8) * 8;
  hipLaunchKernelGGL((k_gemm3m<T, TMT, TNT, WR, WC, KC>), dim3((unsigned)nwg), dim3(WR * WC * 64), 0, s, nmat, ntc,
                     ntile, mstride, M, ldm, r0, m, c0, n, k0, K);
}

// tile shape by operand shape: the smallest row/column tile (32, 64, 128) covering m / n, so narrow
// and short updates do not pay for empty MFMA blocks; 128x128 tiles (64x64 per wave) for the large
// Schur updates halve their re-reads of A and B.  complex128 stops at 128x64 (registers).
template <typename T>
void gemm_region(int nmat, long mstride, cplx<T>* M, int ldm, int r0, int m, int c0, int n, int k0, int K, int mode,
                 const cplx<T>* inv, long istride, hipStream_t s, LuTimer* timer) {
  const int tm = m > 64 ? 128 : m > 32 ? 64 : 32;
  const int tn = n > 64 ? 128 : n > 32 ? 64 : 32;
  // the largest tile shape of this precision is the one the timer brackets with events
  const bool timed = timer && tm == 128 && tn == 128;
  if (timed) {
    timer->mark(s);
    timer->flops += 8.0 * m * (double)n * K * nmat;
  }
#define ADGA_G3(TMT, TNT, WR, WC, KC)                                                                         \
  gemm3m_launch<T, TMT, TNT, WR, WC, KC>(nmat, mstride, M, ldm, r0, m, c0, n, k0, K, s);                     \
  if (timed) timer->mark(s);                                                                                  \
  return;
  if (mode == 0) {
    if constexpr (sizeof(T) == 4) {
      if (tm == 128 && tn == 128) { ADGA_G3(128, 128, 4, 2, 32) }
      if (tm == 128 && tn == 64) { ADGA_G3(128, 64, 4, 2, 16) }
      if (tm == 64 && tn == 128) { ADGA_G3(64, 128, 2, 4, 16) }
      if (tm == 32 && tn == 128) { ADGA_G3(32, 128, 1, 4, 16) }
    } else {
      if (tm == 128 && tn >= 64) { ADGA_G3(128, 64, 4, 2, 16) }
    }
  }
#undef ADGA_G3
#define ADGA_G(TMT, TNT, WR, WC) \
  gemm_launch<T, TMT, TNT, WR, WC>(nmat, mstride, M, ldm, r0, m, c0, n, k0, K, mode, inv, istride, s)
  if (tm == 128 && tn == 128) {
    if constexpr (sizeof(T) == 4) ADGA_G(128, 128, 2, 2);
    else ADGA_G(128, 64, 2, 2);
  } else if (tm == 128 && tn == 64) {
    ADGA_G(128, 64, 2, 2);
  } else if (tm == 128) {
    ADGA_G(128, 32, 4, 1);
  } else if (tm == 64 && tn == 128) {
    if constexpr (sizeof(T) == 4) ADGA_G(64, 128, 2, 2);
    else ADGA_G(64, 64, 2, 2);
  } else if (tm == 32 && tn == 128) {
    ADGA_G(32, 128, 1, 4);
  } else if (tm == 32 && tn == 32) {
    ADGA_G(32, 32, 2, 2);
  } else {
    ADGA_G(64, 64, 2, 2);
  }
#undef ADGA_G
  if (timed) timer->mark(s);
}

// An operation is split into its full large tiles and the remainder strips, each launched with the
// tile shape that fits it: at N = 1080 the top-level update is 536 x 536, which 128x128 tiles alone
// would cover with 30% of their MFMAs on padding.  The triangular products (mode 1/2) keep all
// their rows in one tile (in-place), so only their columns are split.
template <typename T>
void gemm_op(int nmat, long mstride, cplx<T>* M, int ldm, int r0, int m, int c0, int n, int k0, int K, int mode,
             const cplx<T>* inv, long istride, hipStream_t s, LuTimer* timer) {
  constexpr int BR = 128, BC = sizeof(T) == 4 ? 128 : 64;  // largest row / column tile
  const int mr = (mode == 0 && m > BR) ? m / BR * BR : m;
  const int nr = n > BC ? n / BC * BC : n;
  const int rs[2][2] = {{r0, mr}, {r0 + mr, m - mr}};
  const int cs[2][2] = {{c0, nr}, {c0 + nr, n - nr}};
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      if (rs[i][1] > 0 && cs[j][1] > 0)
        gemm_region<T>(nmat, mstride, M, ldm, rs[i][0], rs[i][1], cs[j][0], cs[j][1], k0, K, mode, inv, istride, s,
                       timer);
}

}  // namespace

// The factorisation: the host walks the operation list and launches one batched kernel per
// operation (panels: a workgroup per matrix; GEMM/TRSM: a workgroup per output tile per matrix).
namespace {
template <typename T>
bool lu_op(int nmat, int N, int W, int ldm, const int* op, cplx<T>* M, int* piv, cplx<T>* inv, int* info,
           hipStream_t s, LuTimer* timer) {
  const long mstride = (long)N * ldm, istride = (long)lu_inv_elems(N);
  if (op[0] == OP_PANEL) {
    switch (panel_cfg(N - op[1])) {
#define ADGA_P(R, NTP)                                                                                      \
  case (NTP == 512 ? 100 : 0) + R:                                                                          \
    hipLaunchKernelGGL((k_panel<T, R, LeafW<T, R>::value, NTP>), dim3(nmat), dim3(NTP), 0, s, N, W, ldm,    \
                       mstride, op[1], op[2], M, piv, inv, istride, info);                                  \
    return true;
      ADGA_P(1, 256) ADGA_P(2, 256) ADGA_P(3, 256) ADGA_P(4, 256) ADGA_P(5, 256) ADGA_P(6, 256)
      ADGA_P(4, 512) ADGA_P(5, 512) ADGA_P(6, 512) ADGA_P(8, 512)
#undef ADGA_P
      default: return false;
    }
  } else if (op[0] == OP_TRSM_L || op[0] == OP_TRSM_U) {
    const int K = op[2] - op[1];
    gemm_op<T>(nmat, mstride, M, ldm, 0, K, op[3], op[4] - op[3], op[1], K, op[0] == OP_TRSM_L ? 1 : 2, inv, istride,
               s, nullptr);
  } else {
    gemm_op<T>(nmat, mstride, M, ldm, op[1], op[2] - op[1], op[3], op[4] - op[3], op[5], op[6] - op[5], 0, inv,
               istride, s, timer);
  }
  return true;
}
}  // namespace

// With a second stream the batch is split in two halves, the second one operation behind the
// first on its own stream: while one half runs a latency-bound panel or a small update, the other
// half's operation can use the rest of the machine.  The large Schur updates (those with 128x128
// tiles, the dominant kernel) run on the whole batch on the first stream, between a join and a
// fork, so they own the machine; after each, the second stream starts one operation behind again.
// `lag` and `join` are events the caller owns.
template <typename T>
bool launch_lu(int nmat, int N, int W, int ldm, const int* ops, size_t nops, cplx<T>* M, int* piv, cplx<T>* inv,
               int* info, hipStream_t s, LuTimer* timer, const LuLanes* lanes) {
  if (panel_cfg(N) < 0 || N > PERM_MAX_N || nmat < 1) return false;
  const long mstride = (long)N * ldm, istride = (long)lu_inv_elems(N);
  // the batch in nl parts of about equal size (at least 8 matrices each), part j on lane j; after each join
  // lane j starts its next operation when lane j-1 has finished it, so the parts run staggered by one
  // operation and one part's latency-bound panel or thin update runs beside another's; the large Schur
  // updates run on the whole batch on s alone, between a join of every lane and the next staggered start
  const int nl = (lanes && lanes->n > 1) ? std::min(std::min(lanes->n, LU_MAX_LANES), std::max(1, nmat / 8)) : 1;
  hipStream_t st[LU_MAX_LANES];
  int m0[LU_MAX_LANES + 1];
  for (int j = 0; j < nl; ++j) {
    st[j] = j == 0 ? s : lanes->s[j];
    m0[j] = (int)((long)nmat * j / nl);
  }
  m0[nl] = nmat;
  auto part_op = [&](int j, const int* op, LuTimer* t) {
    const int a = m0[j], cnt = m0[j + 1] - a;
    return lu_op<T>(cnt, N, W, ldm, op, M + (long)a * mstride, piv + (long)a * N, inv + (long)a * istride,
                    info + a, st[j], t);
  };
  auto sync = [](hipEvent_t e, hipStream_t from, hipStream_t to) {
    return hipEventRecord(e, from) == hipSuccess && hipStreamWaitEvent(to, e, 0) == hipSuccess;
  };
  hipLaunchKernelGGL(k_lu_init, dim3((nmat + 255) / 256), dim3(256), 0, s, nmat, info);
  bool uinv = false;
  auto launch_uinv = [&]() {
    const int np = (N + P - 1) / P;
    for (int j = 0; j < nl; ++j)
      hipLaunchKernelGGL(k_lu_uinv<T>, dim3(np, m0[j + 1] - m0[j]), dim3(64), 0, st[j], N, ldm, mstride, istride,
                         M + (long)m0[j] * mstride, inv + (long)m0[j] * istride);
    uinv = true;
  };
  bool relag = true;  // the next operation starts staggered (lane j after lane j-1)
  for (size_t o = 0; o + 8 <= nops; o += 8) {
    const int* op = &ops[o];
    if (op[0] == OP_TRSM_U && !uinv) launch_uinv();  // the upper triangular products need U11^-1
    const bool big = op[0] == OP_GEMM && op[2] - op[1] > 64 && op[4] - op[3] > 64;
    if (nl > 1 && big) {
      for (int j = 1; j < nl; ++j)
        if (!sync(lanes->join[j], st[j], s)) return false;
      if (!lu_op<T>(nmat, N, W, ldm, op, M, piv, inv, info, s, timer)) return false;
      relag = true;
      continue;
    }
    for (int j = 0; j < nl; ++j) {
      if (j > 0 && relag && !sync(lanes->lag[j - 1], st[j - 1], st[j])) return false;
      if (!part_op(j, op, j == 0 ? timer : nullptr)) return false;
    }
    relag = false;
  }
  if (!uinv) launch_uinv();
  for (int j = 0; j < nl; ++j)
    hipLaunchKernelGGL(k_lu_perm, dim3(m0[j + 1] - m0[j]), dim3(NT), 0, st[j], N, piv + (long)m0[j] * N);
  for (int j = 1; j < nl; ++j)
    if (!sync(lanes->join[j], st[j], s)) return false;
  return true;
}
template bool launch_lu<float>(int, int, int, int, const int*, size_t, c32*, int*, c32*, int*, hipStream_t,
                               LuTimer*, const LuLanes*);
template bool launch_lu<double>(int, int, int, int, const int*, size_t, z64*, int*, z64*, int*, hipStream_t,
                                LuTimer*, const LuLanes*);

void LuTimer::mark(hipStream_t s) {
  if (used == ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    ev.push_back(e);
  }
  (void)hipEventRecord(ev[used++], s);
}

double LuTimer::collect() {
  double ms = 0.0;
  for (size_t i = 0; i + 1 < used; i += 2) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) == hipSuccess) ms += t;
  }
  launches = (int)(used / 2);
  used = 0;
  return ms;
}

LuTimer::~LuTimer() {
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}

// columns per group: one group when a solution of nrhs columns (padded to the kernel's column count)
// fits the LDS at two workgroups per CU; else one group on the whole LDS (one workgroup per CU) when it
// fits there; else, with enough matrices to fill the GPU (>= 64), the whole LDS per workgroup when that
// needs fewer groups (each group streams all the factors); else groups as wide as fit at two workgroups
// per CU (few matrices: more groups, more workgroups in flight).  False when not even one column fits.
constexpr int SOLVE_NT = 512;

template <typename T>
bool launch_lu_solve(int nmat, int N, int nrhs, int ldm, const cplx<T>* M, const int* perm, const cplx<T>* inv,
                     const z64* R, z64* Y, bool accumulate, hipStream_t s, const int* active, bool from_m) {
  static const int cgs[] = {1, 2, 4, 9, 16};
  auto pick = [&](int ldsb, bool whole) {
    const int cap = ldsb / (int)sizeof(cplx<T>) / (N + P);  // columns that fit, with P zero rows
    for (int c : cgs)
      if (c <= cap && c >= nrhs) return c;
    int cg = 0;
    if (!whole)
      for (int c : cgs)
        if (c <= cap) cg = c;
    return cg;
  };
  int cg = pick(SOLVE_LDS_BYTES, true);
  bool wide = false;
  if (!cg) {
    const int narrow = pick(SOLVE_LDS_BYTES, false), wcg = pick(SOLVE_LDS_BYTES_WIDE, false);
    const int gn = narrow ? (nrhs + narrow - 1) / narrow : 1 << 30, gw = wcg ? (nrhs + wcg - 1) / wcg : 1 << 30;
    wide = gw < gn && (wcg == 2 || wcg == 4 || wcg == 9 || wcg == 16) && (gw == 1 || nmat >= 64 || !narrow);
    cg = wide ? wcg : narrow;
  }
  if (!cg) return false;
  const int ns = std::min(nrhs, cg), ng = (nrhs + ns - 1) / ns;
  // one workgroup per (matrix, group); with several groups and enough matrices to fill the XCDs, the
  // matrices are dealt round-robin over the 8 XCDs and each matrix's groups take consecutive slots of its
  // XCD (k_lu_solve's mapping; the grid rounded up to whole rounds of 8)
  const int xcd = (ng > 1 && nmat >= 64) ? 1 : 0;
  dim3 grid((unsigned)(xcd ? 8 * ((nmat + 7) / 8) * ng : nmat * ng));
  // 512 threads per workgroup (8 waves of 4 rows): twice the loads in flight per CU of 256 threads (4 waves
  // of 8 rows) under the same LDS limit of workgroups per CU; measured +0.7% points/s refined and +1.5% FP64
  // at N = 1080, +2% (N = 4000) to +6% (N = 2000) with column groups
#define ADGA_SOLVE_CASE(CGV, LDSB)                                                                              \
  if (cg == CGV) {                                                                                              \
    hipLaunchKernelGGL((k_lu_solve<T, CGV, LDSB, SOLVE_NT>), grid, dim3(SOLVE_NT), 0, s, N, nrhs, ns, ng, nmat, \
                       xcd, ldm, M, perm, inv, R, Y, accumulate ? 1 : 0, active, from_m ? 1 : 0);               \
    return true;                                                                                                \
  }
  if (wide) {
    ADGA_SOLVE_CASE(2, SOLVE_LDS_BYTES_WIDE)
    ADGA_SOLVE_CASE(4, SOLVE_LDS_BYTES_WIDE)
    ADGA_SOLVE_CASE(9, SOLVE_LDS_BYTES_WIDE)
    ADGA_SOLVE_CASE(16, SOLVE_LDS_BYTES_WIDE)
    return false;
  }
  ADGA_SOLVE_CASE(1, SOLVE_LDS_BYTES)
  ADGA_SOLVE_CASE(2, SOLVE_LDS_BYTES)
  ADGA_SOLVE_CASE(4, SOLVE_LDS_BYTES)
  ADGA_SOLVE_CASE(9, SOLVE_LDS_BYTES)
  ADGA_SOLVE_CASE(16, SOLVE_LDS_BYTES)
#undef ADGA_SOLVE_CASE
  return false;
}
// whether every nrhs columns of a solution fit one workgroup's LDS (the whole LDS) in complex elements of
// `elem` bytes: then a solve streams the factors once; else once per column group
bool lu_solve_one_group(int N, int nrhs, size_t elem) {
  static const int cgs[] = {1, 2, 4, 9, 16};
  for (int c : cgs)
    if (c >= nrhs && (size_t)c * (N + P) * elem <= (size_t)SOLVE_LDS_BYTES_WIDE) return true;
  return false;
}

template bool launch_lu_solve<float>(int, int, int, int, const c32*, const int*, const c32*, const z64*, z64*, bool,
                                     hipStream_t, const int*, bool);
template bool launch_lu_solve<double>(int, int, int, int, const z64*, const int*, const z64*, const z64*, z64*, bool,
                                      hipStream_t, const int*, bool);

// M[m][r][N + s] = B[m][r][s] (in the factor precision): the right-hand sides appended to each matrix, for
// the factorisation to carry along
template <typename T>
__global__ void k_rhs_into_m(int nmat, int N, int nrhs, int ldm, const z64* __restrict__ B, cplx<T>* __restrict__ M) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)N * nrhs;
  if (idx >= per * nmat) return;
  const long m = idx / per;
  const int r = (int)((idx % per) / nrhs), sidx = (int)(idx % nrhs);
  M[m * N * ldm + (long)r * ldm + N + sidx] = ccast<T>(B[idx]);
}

template <typename T>
void launch_rhs_into_m(int nmat, int N, int nrhs, int ldm, const z64* B, cplx<T>* M, hipStream_t s) {
  const long tot = (long)nmat * N * nrhs;
  hipLaunchKernelGGL(k_rhs_into_m<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, nmat, N, nrhs, ldm, B, M);
}
template void launch_rhs_into_m<float>(int, int, int, int, const z64*, c32*, hipStream_t);
template void launch_rhs_into_m<double>(int, int, int, int, const z64*, z64*, hipStream_t);

}  // namespace adga
