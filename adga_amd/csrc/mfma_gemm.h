// Complex GEMM tiles on gfx950 MFMA (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64) for the LU's
// Schur updates and triangular products and the FP64 BSE residual products.  gemm_tile forms a
// complex product as a real product with K doubled and re/im interleaved along it (the lane's K
// slot selecting (re, -im) or (im, re) of the A element and (re, im) of the B element);
// gemm3m_tile forms it from three real products (the 3M method), with C reached through accessor
// functors so that the same tile serves row-major matrices and the batched point layouts.
#pragma once

#include "adga_common.h"

namespace adga {

constexpr int GEMM_NT = 256;  // threads per workgroup (4 waves)

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

// ---------------------------------------------------------------- row-major tile (the LU's GEMMs)
// C[0:m, 0:n] <- C - A B (assign = false) or A B (assign = true) for row-major A, B, C (the LU's
// updates and triangular products), the four waves in a WR x WC grid, each a (TMT/WR) x (TNT/WC)
// block of 16x16 MFMA tiles; K staged through LDS in chunks of KC, the next chunk's global loads
// issued into registers before the current chunk's MFMAs; the output is written after all of K
// has been read by every wave, so C may alias B when the tile spans all K rows of B (in-place
// triangular products: m = K <= TMT).  Staging:
// 16-byte global loads (two complex64 or one complex128 per load) without bounds checks on
// interior tiles of aligned operands, int offsets, and LDS layouts read without bank conflicts by
// the MFMA fragments: A as As[row][k] (row stride KC complex + 16 bytes), B as Bs[k][col] (row
// stride TNT complex + 128 bytes, so the lanes of the two K slots fall in the two bank halves).
template <typename T, int TMT, int TNT>
struct RmCfg {
  static constexpr int CS = (int)sizeof(cplx<T>);
  static constexpr int VEC = 16 / CS;  // complex per 16-byte load
  static constexpr int budget = 36864;
  static constexpr int per_k = TMT * CS + (TNT * CS + 128);  // A column + B row bytes per k
  static constexpr int KC0 = (budget - TMT * 16) / per_k;
  static constexpr int KC = KC0 >= 32 ? 32 : KC0 >= 16 ? 16 : 8;  // keeps the per-thread load counts whole
  static constexpr int LDA = KC + 16 / CS;    // As row stride (complex)
  static constexpr int LDB = TNT + 128 / CS;  // Bs row stride (complex)
  static constexpr int LDS = (TMT * LDA + KC * LDB) * CS;
};

template <typename T, int TMT, int TNT, int WR, int WC>
__device__ __forceinline__ void gemm_tile(const cplx<T>* __restrict__ A, int lda, const cplx<T>* B, int ldb,
                                          cplx<T>* Cm, int ldc, int m, int n, int K, bool assign, int rt, int ct,
                                          unsigned char* lds) {
  using V = typename Mfma<T>::V;
  using C = cplx<T>;
  using Cfg = RmCfg<T, TMT, TNT>;
  static_assert(WR * WC == GEMM_NT / 64, "four waves");
  constexpr int WTM = TMT / WR, WTN = TNT / WC, RB = WTM / 16, CB = WTN / 16;
  static_assert(RB >= 1 && CB >= 1 && WTM % 16 == 0 && WTN % 16 == 0, "wave tile of 16x16 blocks");
  constexpr int KC = Cfg::KC, VEC = Cfg::VEC, LDA = Cfg::LDA, LDB = Cfg::LDB;
  constexpr int AV = KC / VEC, BV = TNT / VEC;  // 16-byte vectors per A row / B row of a chunk
  constexpr int PA = TMT * AV / GEMM_NT, PB = KC * BV / GEMM_NT;
  static_assert(KC >= 8 && KC % VEC == 0 && (TMT * AV) % GEMM_NT == 0 && (KC * BV) % GEMM_NT == 0, "K chunk");
  struct alignas(16) CV { C v[VEC]; };
  C* As = reinterpret_cast<C*>(lds);
  C* Bs = As + TMT * LDA;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wr = (wave / WC) * WTM, wc = (wave % WC) * WTN;
  const int li = lane & 15, lg = lane >> 4, part = lg & 1, kcl = lg >> 1;
  const T sgn = assign ? (T)1 : (T)-1;
  // unchecked 16-byte loads need every row start 16-byte aligned and the whole tile inside
  const bool aligned = ((reinterpret_cast<size_t>(A) | reinterpret_cast<size_t>(B)) % 16 == 0) &&
                       (lda % VEC == 0) && (ldb % VEC == 0);
  const bool inside = aligned && rt + TMT <= m && ct + TNT <= n;
  V accr[RB][CB], acci[RB][CB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int row = rt + wr + rb * 16 + Mfma<T>::row(lg, v), col = ct + wc + cb * 16 + li;
        C c = (!assign && row < m && col < n) ? Cm[row * ldc + col] : mk<T>(0, 0);
        accr[rb][cb][v] = c.re;
        acci[rb][cb][v] = c.im;
      }
  CV pa[PA], pb[PB];
  auto fetch = [&](int kb) {
    if (inside && kb + KC <= K) {
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        const int e = t + i * GEMM_NT, row = e / AV, k = (e % AV) * VEC;
        pa[i] = *reinterpret_cast<const CV*>(A + (rt + row) * lda + kb + k);
      }
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        const int e = t + i * GEMM_NT, k = e / BV, col = (e % BV) * VEC;
        pb[i] = *reinterpret_cast<const CV*>(B + (kb + k) * ldb + ct + col);
      }
    } else {
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        const int e = t + i * GEMM_NT, row = rt + e / AV, k = kb + (e % AV) * VEC;
#pragma unroll
        for (int j = 0; j < VEC; ++j) pa[i].v[j] = (row < m && k + j < K) ? A[row * lda + k + j] : mk<T>(0, 0);
      }
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        const int e = t + i * GEMM_NT, k = kb + e / BV, col = ct + (e % BV) * VEC;
#pragma unroll
        for (int j = 0; j < VEC; ++j) pb[i].v[j] = (k < K && col + j < n) ? B[k * ldb + col + j] : mk<T>(0, 0);
      }
    }
  };
  fetch(0);
  for (int kb = 0; kb < K; kb += KC) {
    if (kb) __syncthreads();  // the previous chunk's MFMAs are done with LDS
#pragma unroll
    for (int i = 0; i < PA; ++i) {
      const int e = t + i * GEMM_NT, row = e / AV, k = (e % AV) * VEC;
      *reinterpret_cast<CV*>(As + row * LDA + k) = pa[i];
    }
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      const int e = t + i * GEMM_NT, k = e / BV, col = (e % BV) * VEC;
      *reinterpret_cast<CV*>(Bs + k * LDB + col) = pb[i];
    }
    __syncthreads();
    const int kn = kb + KC;
    if (kn < K) fetch(kn);  // in flight during this chunk's MFMAs
    const int kend = min(KC, K - kb);
    const C* Ap = As + (wr + li) * LDA + kcl;
    const C* Bp = Bs + kcl * LDB + wc + li;
#pragma unroll 2
    for (int kk = 0; kk < kend; kk += 2) {
      T ar[RB], ai[RB], bb[CB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        const C a = Ap[rb * 16 * LDA + kk];
        ar[rb] = sgn * (part ? -a.im : a.re);
        ai[rb] = sgn * (part ? a.re : a.im);
      }
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const C b = Bp[kk * LDB + cb * 16];
        bb[cb] = part ? b.im : b.re;
      }
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
          accr[rb][cb] = Mfma<T>::op(ar[rb], bb[cb], accr[rb][cb]);
          acci[rb][cb] = Mfma<T>::op(ai[rb], bb[cb], acci[rb][cb]);
        }
    }
  }
  __syncthreads();  // every wave has consumed its operands before the output (may alias B) lands
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int row = rt + wr + rb * 16 + Mfma<T>::row(lg, v), col = ct + wc + cb * 16 + li;
        if (row < m && col < n) Cm[row * ldc + col] = mk<T>(accr[rb][cb][v], acci[rb][cb][v]);
      }
}


// ---------------------------------------------------------------- 3M tile (three real products)
// C[0:m, 0:n] <- C -/+ A B (assign = false, sub = true/false) or -/+ A B (assign = true) for
// row-major A and B, C read through cload(row, col) and written through cstore(row, col, value),
// with the complex product formed from three real products on the MFMA instead of four:
// X = Ar Br, Y = Ai Bi, Z = (Ar + Ai)(Br + Bi), Re(AB) = X - Y, Im(AB) = Z - X - Y (the 3M method:
// normwise stable, the imaginary part's error bounded by u (|Ar| + |Ai|)(|Br| + |Bi|) instead of
// u |A||B|; N. J. Higham, SIAM J. Matrix Anal. Appl. 13 (1992) 681-687).  Each MFMA takes four
// consecutive k, lane l supplying k = l/16.  WR x WC waves (the workgroup is WR * WC * 64
// threads); two LDS buffers, chunk kb + KC staged into one while chunk kb is multiplied out of the
// other, so each chunk costs one barrier.  With assign = false, C is read during the last chunk,
// so it must not alias A or B; with assign = true, C is written after the last barrier and may
// alias B when the tile spans all K rows of B (m = K <= TMT).
template <typename T, int TMT, int TNT, int WR, int WC, int KC>
struct G3Cfg {
  static constexpr int NTH = WR * WC * 64;
  static constexpr int CS = (int)sizeof(cplx<T>);
  static constexpr int VEC = 16 / CS;        // complex per 16-byte load
  static constexpr int LDA = KC + 16 / CS;   // As row stride (complex): the 16 lanes of a K slot hit distinct banks
  static constexpr int LDB = TNT + 128 / CS;  // Bs row stride (complex): K slots 0/1 in opposite bank halves
  static constexpr int BUF = TMT * LDA + KC * LDB;  // complex per buffer
  static constexpr int LDS = 2 * BUF * CS;
};

template <typename T, int TMT, int TNT, int WR, int WC, int KC, class CL, class CS>
__device__ __forceinline__ void gemm3m_tile(const cplx<T>* __restrict__ A, int lda, const cplx<T>* B, int ldb,
                                            CL cload, CS cstore, int m, int n, int K, bool assign, bool sub,
                                            int rt, int ct, unsigned char* lds) {
  using V = typename Mfma<T>::V;
  using C = cplx<T>;
  using Cfg = G3Cfg<T, TMT, TNT, WR, WC, KC>;
  constexpr int NTH = Cfg::NTH, VEC = Cfg::VEC, LDA = Cfg::LDA, LDB = Cfg::LDB, BUF = Cfg::BUF;
  constexpr int WTM = TMT / WR, WTN = TNT / WC, RB = WTM / 16, CB = WTN / 16;
  static_assert(RB >= 1 && CB >= 1 && WTM % 16 == 0 && WTN % 16 == 0, "wave tile of 16x16 blocks");
  constexpr int AV = KC / VEC, BV = TNT / VEC;  // 16-byte vectors per A row / B row of a chunk
  constexpr int PA = TMT * AV / NTH, PB = KC * BV / NTH;
  static_assert(KC % 4 == 0 && KC % VEC == 0 && PA >= 1 && PB >= 1 && (TMT * AV) % NTH == 0 && (KC * BV) % NTH == 0,
                "K chunk");
  struct alignas(16) CV { C v[VEC]; };
  C* const L = reinterpret_cast<C*>(lds);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wr = (wave / WC) * WTM, wc = (wave % WC) * WTN;
  const int li = lane & 15, lg = lane >> 4;
  const bool aligned = ((reinterpret_cast<size_t>(A) | reinterpret_cast<size_t>(B)) % 16 == 0) &&
                       (lda % VEC == 0) && (ldb % VEC == 0);
  const bool inside = aligned && rt + TMT <= m && ct + TNT <= n;
  V x[RB][CB], y[RB][CB], z[RB][CB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) x[rb][cb][v] = y[rb][cb][v] = z[rb][cb][v] = (T)0;
  CV pa[PA], pb[PB];
  auto fetch = [&](int kb) {
    if (inside && kb + KC <= K) {
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        const int e = t + i * NTH, row = e / AV, k = (e % AV) * VEC;
        pa[i] = *reinterpret_cast<const CV*>(A + (rt + row) * lda + kb + k);
      }
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        const int e = t + i * NTH, k = e / BV, col = (e % BV) * VEC;
        pb[i] = *reinterpret_cast<const CV*>(B + (kb + k) * ldb + ct + col);
      }
    } else {
#pragma unroll
      for (int i = 0; i < PA; ++i) {
        const int e = t + i * NTH, row = rt + e / AV, k = kb + (e % AV) * VEC;
#pragma unroll
        for (int j = 0; j < VEC; ++j) pa[i].v[j] = (row < m && k + j < K) ? A[row * lda + k + j] : mk<T>(0, 0);
      }
#pragma unroll
      for (int i = 0; i < PB; ++i) {
        const int e = t + i * NTH, k = kb + e / BV, col = ct + (e % BV) * VEC;
#pragma unroll
        for (int j = 0; j < VEC; ++j) pb[i].v[j] = (k < K && col + j < n) ? B[k * ldb + col + j] : mk<T>(0, 0);
      }
    }
  };
  auto stage = [&](int b) {
    C* As = L + b * BUF;
    C* Bs = As + TMT * LDA;
#pragma unroll
    for (int i = 0; i < PA; ++i) {
      const int e = t + i * NTH, row = e / AV, k = (e % AV) * VEC;
      *reinterpret_cast<CV*>(As + row * LDA + k) = pa[i];
    }
#pragma unroll
    for (int i = 0; i < PB; ++i) {
      const int e = t + i * NTH, k = e / BV, col = (e % BV) * VEC;
      *reinterpret_cast<CV*>(Bs + k * LDB + col) = pb[i];
    }
  };
  fetch(0);
  stage(0);
  if (KC < K) fetch(KC);
  __syncthreads();
  auto compute = [&](int b, int kend) {  // k in [kend, KC) of a last, partial chunk are staged zeros
    const C* Ap = L + b * BUF + (wr + li) * LDA + lg;
    const C* Bp = L + b * BUF + TMT * LDA + lg * LDB + wc + li;
#pragma unroll 2
    for (int kk = 0; kk < kend; kk += 4) {
      T ar[RB], ai[RB], as[RB], br[CB], bi[CB], bs[CB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) {
        const C a = Ap[rb * 16 * LDA + kk];
        ar[rb] = a.re;
        ai[rb] = a.im;
        as[rb] = a.re + a.im;
      }
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const C bv = Bp[kk * LDB + cb * 16];
        br[cb] = bv.re;
        bi[cb] = bv.im;
        bs[cb] = bv.re + bv.im;
      }
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
          x[rb][cb] = Mfma<T>::op(ar[rb], br[cb], x[rb][cb]);
          y[rb][cb] = Mfma<T>::op(ai[rb], bi[cb], y[rb][cb]);
          z[rb][cb] = Mfma<T>::op(as[rb], bs[cb], z[rb][cb]);
        }
    }
  };
  int b = 0, kb = 0;
  for (; kb + KC < K; kb += KC) {
    compute(b, KC);
    // the other buffer was last read in the previous chunk, before the barrier that ended it
    stage(b ^ 1);
    if (kb + 2 * KC < K) fetch(kb + 2 * KC);
    __syncthreads();
    b ^= 1;
  }
  // the last chunk (no fetch in flight): C of the update form is loaded beside its products
  C cin[RB][CB][4];
  if (!assign) {
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = rt + wr + rb * 16 + Mfma<T>::row(lg, v), col = ct + wc + cb * 16 + li;
          cin[rb][cb][v] = (row < m && col < n) ? cload(row, col) : mk<T>(0, 0);
        }
  }
  if (kb < K) compute(b, K - kb);
  __syncthreads();
  // the last barrier has passed: every wave's reads of A and B are done, so C may alias B
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = rt + wr + rb * 16 + Mfma<T>::row(lg, v), col = ct + wc + cb * 16 + li;
        if (row < m && col < n) {
          T re = x[rb][cb][v] - y[rb][cb][v], im = (z[rb][cb][v] - x[rb][cb][v]) - y[rb][cb][v];
          if (sub) {
            re = -re;
            im = -im;
          }
          if (!assign) {
            re = cin[rb][cb][v].re + re;
            im = cin[rb][cb][v].im + im;
          }
          cstore(row, col, mk<T>(re, im));
        }
      }
}

}  // namespace adga
