// The non-local interaction V(q) on gfx950: the density-channel BSE term (main.F90:710-719), the V
// terms of the dynamic EOM's m_tot (eom_module.f90:183-207) and the non-local Hartree and Fock terms
// of calc_eom_static (eom_module.f90:72-108).
//
// With V the density BSE matrix gains a term of rank n2:
//   A_d = 1 - blockdiag(S) C_d - W (1 + gamma^w_d),   W[d*n2 + a][s] = (2/beta^2) (chi0^q(nu_d) V(q))(a, s),
// so, in the transposed system M = A^T that the LU factors (M eta^T = B),
//   M_d   -= (1 + gamma^w_d)^T W^T                          (k_vq_build)
//   B_d   += (1 + gamma^w_d)^T (W^T (1 + gamma^w_d)^T)      (the right-hand side of eta A = (1+gamma)(1-A))
//   R_d   += (1 + gamma^w_d)^T (W^T Y_d)                     (the residual B - M Y)
// The two products with W^T are n2 x n2 per point (k_vq_small); k_vq_apply adds (1+gamma)^T times them.
#include "adga_kernels.h"

namespace adga {

// W[b][x][s] = (2/beta^2) sum_c chi0q_b(a, c, d) V_{qpos[b]}(c, s), x = d*n2 + a (main.F90:712-713).
__global__ void k_vq_w(Geom g, int npts, const int* __restrict__ qpos, const z64* __restrict__ vq,
                       const z64* __restrict__ chi0q, z64* __restrict__ W) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npts * g.N * g.n2) return;
  const int s = (int)(idx % g.n2);
  const long bx = idx / g.n2;
  const int x = (int)(bx % g.N), b = (int)(bx / g.N);
  const int d = x / g.n2, a = x % g.n2;
  const z64* cq = chi0q + ((long)b * g.nf + d) * g.n4;  // (a, c) at a + c*n2
  const z64* v = vq + (long)qpos[b] * g.n4;            // (c, s) at c + s*n2
  z64 acc = mk(0.0, 0.0);
  for (int c = 0; c < g.n2; ++c) acc = cfma(cq[a + c * g.n2], v[c + s * g.n2], acc);
  W[idx] = cscale(acc, 2.0 / (g.beta * g.beta));
}

// M_{2b}(r, x) -= sum_s W_b[x][s] opg_d(s, r): the density matrix of point b (rows r of M = A^T).
template <typename T>
__global__ void k_vq_build(Geom g, int npts, int ldm, const z64* __restrict__ W, const z64* __restrict__ opgd,
                           cplx<T>* __restrict__ M) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long NN = (long)g.N * g.N;
  if (idx >= (long)npts * NN) return;
  const int x = (int)(idx % g.N);
  const int r = (int)((idx / g.N) % g.N);
  const int b = (int)(idx / NN);
  const z64* w = W + ((long)b * g.N + x) * g.n2;
  const z64* o = opgd + (long)r * g.n2;
  z64 acc = mk(0.0, 0.0);
  for (int s = 0; s < g.n2; ++s) acc = cfma(w[s], o[s], acc);
  cplx<T>* m = M + ((long)(2 * b) * g.N + r) * ldm + x;
  const z64 v = csub(ccast<double>(*m), acc);
  *m = ccast<T>(v);
}

// Q_b(s, s') = sum_x W_b[x][s] X_b(x, s') with X = (1 + gamma^w_d)^T (Y null) or the density solution
// Y_{2b}[x][s']; Q (n2 x n2 per point) in Fortran order s + s'*n2.  One wave per output, lanes over x.
__global__ void __launch_bounds__(64) k_vq_small(Geom g, const z64* __restrict__ W, const z64* __restrict__ opgd,
                                                 const z64* __restrict__ Y, z64* __restrict__ Q) {
  const int e = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
  const int s = e % g.n2, sp = e / g.n2;
  const z64* w = W + (long)b * g.N * g.n2;
  double re = 0.0, im = 0.0;
  for (int x = lane; x < g.N; x += 64) {
    const z64 xv = Y ? Y[((long)(2 * b) * g.N + x) * g.n2 + sp] : opgd[sp + (long)x * g.n2];
    const z64 p = cmul(w[(long)x * g.n2 + s], xv);
    re += p.re;
    im += p.im;
  }
  for (int o = 32; o > 0; o >>= 1) {
    re += __shfl_down(re, o, 64);
    im += __shfl_down(im, o, 64);
  }
  if (lane == 0) Q[(long)b * g.n4 + e] = mk(re, im);
}

// Out_{2b}[r][s'] += sum_s opg_d(s, r) Q_b(s, s') (right-hand side / residual layout [(2b+ch) N + r] n2 + s').
__global__ void k_vq_apply(Geom g, int npts, const z64* __restrict__ opgd, const z64* __restrict__ Q,
                           z64* __restrict__ Out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npts * g.N * g.n2) return;
  const int sp = (int)(idx % g.n2);
  const long br = idx / g.n2;
  const int r = (int)(br % g.N), b = (int)(br / g.N);
  const z64* o = opgd + (long)r * g.n2;
  const z64* q = Q + (long)b * g.n4 + (long)sp * g.n2;
  z64 acc = mk(0.0, 0.0);
  for (int s = 0; s < g.n2; ++s) acc = cfma(o[s], q[s], acc);
  z64* out = Out + ((long)(2 * b) * g.N + r) * g.n2 + sp;
  *out = cadd(*out, acc);
}

void launch_vq_w(const Geom& g, int npts, const int* qpos, const z64* vq, const z64* chi0q, z64* W, hipStream_t s) {
  const long tot = (long)npts * g.N * g.n2;
  if (tot > 0)
    hipLaunchKernelGGL(k_vq_w, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, qpos, vq, chi0q, W);
}

template <typename T>
void launch_vq_build(const Geom& g, int npts, int ldm, const VqTerm& v, cplx<T>* M, hipStream_t s) {
  const long tot = (long)npts * g.N * g.N;
  if (tot > 0)
    hipLaunchKernelGGL((k_vq_build<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, ldm, v.W,
                       v.opgd, M);
}
template void launch_vq_build<float>(const Geom&, int, int, const VqTerm&, c32*, hipStream_t);
template void launch_vq_build<double>(const Geom&, int, int, const VqTerm&, z64*, hipStream_t);

void launch_vq_add(const Geom& g, int npts, const VqTerm& v, const z64* Y, z64* Out, hipStream_t s) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(k_vq_small, dim3(g.n4, npts), dim3(64), 0, s, g, v.W, v.opgd, Y, v.Q);
  const long tot = (long)npts * g.N * g.n2;
  hipLaunchKernelGGL(k_vq_apply, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, v.opgd, v.Q, Out);
}

// Fock term at omega = 0 (eom_module.f90:88-106): sigma_hf(i, l, kk) -= sum_b sum_{j,k'} V_b({i k'}, {l j}) / nqp
// * n_fock(k_eom(kk) - q_b, j, k'), with V(i1, i2) at i1 + i2*n2 and n_fock(nkp, n, n) in Fortran order.
// One thread per (i, l, kk); the batch is summed in a fixed order.
__global__ void k_eom_fock(Geom g, int npts, const int* __restrict__ qpos, const int* __restrict__ qidx,
                           const int* __restrict__ keom, const z64* __restrict__ vq, const z64* __restrict__ nfock,
                           z64* __restrict__ shf) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)g.n2 * g.nkp_eom) return;
  const int il = (int)(idx % g.n2), kk = (int)(idx / g.n2);
  const int i = il % g.n, l = il / g.n;
  const int k = keom[kk] - 1;
  z64 acc = mk(0.0, 0.0);
  for (int b = 0; b < npts; ++b) {
    const int kq = k_minus_q(k, qidx[b], g.nkx, g.nky, g.nkz);
    const z64* v = vq + (long)qpos[b] * g.n4;
    for (int j = 0; j < g.n; ++j)
      for (int kp = 0; kp < g.n; ++kp)
        acc = cfma(v[(i * g.n + kp) + (long)(l * g.n + j) * g.n2], nfock[kq + (long)j * g.nkp + (long)kp * g.nkp * g.n],
                   acc);
  }
  z64* o = shf + i + (long)l * g.n + (long)kk * g.n2;
  *o = csub(*o, cscale(acc, 1.0 / g.nqp));
}

// Non-local Hartree term at q = the first q of the list, omega = 0 (eom_module.f90:72-84):
// sigma_hf(i, i, kk) += 2 sum_j V({ii}, {jj}) n_dmft(j) for every kk.
__global__ void k_eom_hartree_v(Geom g, const z64* __restrict__ v, const z64* __restrict__ ndmft,
                                z64* __restrict__ shf) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)g.n * g.nkp_eom) return;
  const int i = (int)(idx % g.n), kk = (int)(idx / g.n);
  z64 acc = mk(0.0, 0.0);
  for (int j = 0; j < g.n; ++j) acc = cfma(v[(i * g.n + i) + (long)(j * g.n + j) * g.n2], ndmft[j], acc);
  z64* o = shf + i + (long)i * g.n + (long)kk * g.n2;
  *o = cadd(*o, cscale(acc, 2.0));
}

void launch_eom_fock(const Geom& g, int npts, const int* qpos, const int* qidx, const int* keom, const z64* vq,
                     const z64* nfock, z64* shf, hipStream_t s) {
  const long tot = (long)g.n2 * g.nkp_eom;
  if (npts > 0)
    hipLaunchKernelGGL(k_eom_fock, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, qpos, qidx, keom, vq,
                       nfock, shf);
}

void launch_eom_hartree_v(const Geom& g, const z64* v, const z64* ndmft, z64* shf, hipStream_t s) {
  const long tot = (long)g.n * g.nkp_eom;
  hipLaunchKernelGGL(k_eom_hartree_v, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, v, ndmft, shf);
}

}  // namespace adga
