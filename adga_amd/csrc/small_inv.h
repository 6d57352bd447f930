// Closed-form and Gauss-Jordan inverses of small complex matrices held in registers (n <= 8), shared
// by the G(k, nu) table and the one-particle routines.
#pragma once

#include "adga_common.h"

namespace adga {

// ---------------------------------------------------------------- small inverses
// Closed-form adjugate for n = 1, 2, 3 (the reference's choice:
// one_particle_quant_module.f90:245-256 with aux.f90:288-325).
template <int NN>
__device__ __forceinline__ void inv_small(z64 (&a)[NN][NN]);

template <>
__device__ __forceinline__ void inv_small<1>(z64 (&a)[1][1]) {
  a[0][0] = crecip(a[0][0]);
}
template <>
__device__ __forceinline__ void inv_small<2>(z64 (&a)[2][2]) {
  z64 det = csub(cmul(a[0][0], a[1][1]), cmul(a[0][1], a[1][0]));
  z64 d = crecip(det);
  z64 b00 = cmul(d, a[1][1]), b01 = cmul(d, a[0][1]), b10 = cmul(d, a[1][0]), b11 = cmul(d, a[0][0]);
  a[0][0] = b00;
  a[0][1] = mk(-b01.re, -b01.im);
  a[1][0] = mk(-b10.re, -b10.im);
  a[1][1] = b11;
}
template <>
__device__ __forceinline__ void inv_small<3>(z64 (&a)[3][3]) {
  // cofactors c_ij = (-1)^(i+j) minor_ij; inverse = adj / det with adj = c^T
  z64 c00 = csub(cmul(a[1][1], a[2][2]), cmul(a[1][2], a[2][1]));
  z64 c01 = csub(cmul(a[1][2], a[2][0]), cmul(a[1][0], a[2][2]));
  z64 c02 = csub(cmul(a[1][0], a[2][1]), cmul(a[1][1], a[2][0]));
  z64 c10 = csub(cmul(a[0][2], a[2][1]), cmul(a[0][1], a[2][2]));
  z64 c11 = csub(cmul(a[0][0], a[2][2]), cmul(a[0][2], a[2][0]));
  z64 c12 = csub(cmul(a[0][1], a[2][0]), cmul(a[0][0], a[2][1]));
  z64 c20 = csub(cmul(a[0][1], a[1][2]), cmul(a[0][2], a[1][1]));
  z64 c21 = csub(cmul(a[0][2], a[1][0]), cmul(a[0][0], a[1][2]));
  z64 c22 = csub(cmul(a[0][0], a[1][1]), cmul(a[0][1], a[1][0]));
  z64 det = cadd(cadd(cmul(a[0][0], c00), cmul(a[0][1], c01)), cmul(a[0][2], c02));
  z64 d = crecip(det);
  a[0][0] = cmul(d, c00); a[0][1] = cmul(d, c10); a[0][2] = cmul(d, c20);
  a[1][0] = cmul(d, c01); a[1][1] = cmul(d, c11); a[1][2] = cmul(d, c21);
  a[2][0] = cmul(d, c02); a[2][1] = cmul(d, c12); a[2][2] = cmul(d, c22);
}

// Gauss-Jordan with partial pivoting for 4 <= n <= 8 (the reference calls
// LAPACK zgetrf/zgetri here, lapack_module.f90:40-70).  Fully unrolled so the
// matrix stays in registers.
template <int NN>
__device__ __forceinline__ void inv_small(z64 (&a)[NN][NN]) {
  z64 b[NN][NN];
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int j = 0; j < NN; ++j) b[i][j] = mk(i == j ? 1.0 : 0.0, 0.0);
#pragma unroll
  for (int c = 0; c < NN; ++c) {
    // pivot search in column c among rows >= c; swap by predicated register moves
    double best = -1.0;
    int p = c;
#pragma unroll
    for (int r = c; r < NN; ++r) {
      double v = cabs1(a[r][c]);
      if (v > best) { best = v; p = r; }
    }
#pragma unroll
    for (int r = c + 1; r < NN; ++r) {
      if (r == p) {
#pragma unroll
        for (int j = 0; j < NN; ++j) {
          z64 t = a[c][j]; a[c][j] = a[r][j]; a[r][j] = t;
          t = b[c][j]; b[c][j] = b[r][j]; b[r][j] = t;
        }
      }
    }
    z64 ip = crecip(a[c][c]);
#pragma unroll
    for (int j = 0; j < NN; ++j) { a[c][j] = cmul(a[c][j], ip); b[c][j] = cmul(b[c][j], ip); }
#pragma unroll
    for (int r = 0; r < NN; ++r) {
      if (r == c) continue;
      z64 f = a[r][c];
#pragma unroll
      for (int j = 0; j < NN; ++j) { a[r][j] = cfms(f, a[c][j], a[r][j]); b[r][j] = cfms(f, b[c][j], b[r][j]); }
    }
  }
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int j = 0; j < NN; ++j) a[i][j] = b[i][j];
}

}  // namespace adga
