// Shared device/host helpers for the gfx950 hot-path kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace adga {

template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};
using z64 = cplx<double>;  // complex128
using c32 = cplx<float>;   // complex64

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T r, T i) {
  cplx<T> c;
  c.re = r;
  c.im = i;
  return c;
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return mk<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
  return mk<T>(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) {
  return mk<T>(a.re * s, a.im * s);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cconj(cplx<T> a) {
  return mk<T>(a.re, -a.im);
}
// acc + a*b
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cfma(cplx<T> a, cplx<T> b, cplx<T> acc) {
  acc.re = fma(a.re, b.re, acc.re);
  acc.re = fma(-a.im, b.im, acc.re);
  acc.im = fma(a.re, b.im, acc.im);
  acc.im = fma(a.im, b.re, acc.im);
  return acc;
}
// acc - a*b
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cfms(cplx<T> a, cplx<T> b, cplx<T> acc) {
  acc.re = fma(-a.re, b.re, acc.re);
  acc.re = fma(a.im, b.im, acc.re);
  acc.im = fma(-a.re, b.im, acc.im);
  acc.im = fma(-a.im, b.re, acc.im);
  return acc;
}
// 1/a
template <typename T>
__host__ __device__ __forceinline__ cplx<T> crecip(cplx<T> a) {
  T d = a.re * a.re + a.im * a.im;
  return mk<T>(a.re / d, -a.im / d);
}
// a/b with the textbook formula (b conj)/|b|^2
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cdiv(cplx<T> a, cplx<T> b) {
  T d = b.re * b.re + b.im * b.im;
  return mk<T>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
template <typename T>
__host__ __device__ __forceinline__ T cabs1(cplx<T> a) {
  return fabs(a.re) + fabs(a.im);
}
template <typename D, typename S>
__host__ __device__ __forceinline__ cplx<D> ccast(cplx<S> a) {
  return mk<D>((D)a.re, (D)a.im);
}

// Problem geometry shared by all kernels of one context (plain data, passed by value).
struct Geom {
  int n, n2, n4;        // orbitals, n^2, n^4
  int nf;               // fermionic frequencies of the BSE box of the current omega (2*iwf)
  int iwf;              // half of it (iwfmax_small, or less where the box shrinks with |omega|)
  int iws;              // iwfmax_small: the self-energies' frequency box (offset iws, 2*iws frequencies)
  int iwbmax;           // iwbmax_small
  int N;                // n2*nf, BSE dimension
  int nkx, nky, nkz, nkp;
  int glo, nw;          // G table covers nu in [glo, glo+nw)
  int iwmax;            // 1P data frequencies
  int nqp, nkp_eom;
  double beta, mu;
};

// 0-based index of k - q on the k grid (kq_tools.f90:96-108).
__host__ __device__ __forceinline__ int k_minus_q(int k, int q, int nkx, int nky, int nkz) {
  int kz = k % nkz, ky = (k / nkz) % nky, kx = k / (nky * nkz);
  int qz = q % nkz, qy = (q / nkz) % nky, qx = q / (nky * nkz);
  return (kz - qz + nkz) % nkz + ((ky - qy + nky) % nky) * nkz + ((kx - qx + nkx) % nkx) * nky * nkz;
}

}  // namespace adga
