// Sums of the lambda correction in FP64 on the device.  The physical susceptibility of a point is the sum of
// the reduced susceptibility's entries, chi_phys = sum_{i,j} chi[(i,i), (j,j)] (the entries the reference's
// reduced output keeps, hdf5_module.f90:1479-1495); the corrected one is chi^lambda = chi / (1 + lambda chi)
// = 1/(1/chi + lambda), and the sums over all points are S(lambda) = sum chi^lambda and
// S'(lambda) = -sum (chi^lambda)^2.  The reductions are two-stage with a fixed grid (a partial per block, then
// one block per lambda in a fixed order), so repeated sums are bit-identical.
#include "adga_kernels.h"

namespace adga {

namespace {
constexpr int LAM_THREADS = 256;
}

// out[p] = sum_{i,j} chi[(i*n + i) + (j*n + j)*n2 + p*n4]
__global__ void k_chi_phys(int n, long npts, const z64* __restrict__ chi, z64* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const int n2 = n * n;
  const z64* c = chi + p * (long)n2 * n2;
  z64 acc = mk(0.0, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) acc = cadd(acc, c[(i * n + i) + (long)(j * n + j) * n2]);
  out[p] = acc;
}

void launch_chi_phys(int n, long npts, const z64* chi, z64* out, hipStream_t s) {
  hipLaunchKernelGGL(k_chi_phys, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, n, npts, chi, out);
}

__device__ __forceinline__ void block_reduce2(z64& a, z64& b, z64* red) {
  red[threadIdx.x] = a;
  red[LAM_THREADS + threadIdx.x] = b;
  __syncthreads();
  for (int h = LAM_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + h]);
      red[LAM_THREADS + threadIdx.x] = cadd(red[LAM_THREADS + threadIdx.x], red[LAM_THREADS + threadIdx.x + h]);
    }
    __syncthreads();
  }
  a = red[0];
  b = red[LAM_THREADS];
}

// part[(l * nblk + blk) * 2 + {0, 1}] = block blk's sums of chi^lambda_l and (chi^lambda_l)^2; grid (nblk, nl)
__global__ void __launch_bounds__(LAM_THREADS) k_lambda_partial(long npts, const double* __restrict__ lams,
                                                                const z64* __restrict__ chi, z64* __restrict__ part) {
  __shared__ z64 red[2 * LAM_THREADS];
  const int l = blockIdx.y;
  const double lam = lams[l];
  z64 s1 = mk(0.0, 0.0), s2 = mk(0.0, 0.0);
  for (long p = (long)blockIdx.x * LAM_THREADS + threadIdx.x; p < npts; p += (long)gridDim.x * LAM_THREADS) {
    const z64 c = chi[p];
    const z64 cl = cdiv(c, mk(1.0 + lam * c.re, lam * c.im));
    s1 = cadd(s1, cl);
    s2 = cadd(s2, cmul(cl, cl));
  }
  block_reduce2(s1, s2, red);
  if (threadIdx.x == 0) {
    part[((long)l * gridDim.x + blockIdx.x) * 2] = s1;
    part[((long)l * gridDim.x + blockIdx.x) * 2 + 1] = s2;
  }
}

// out[l * 2] = scale * sum, out[l * 2 + 1] = -scale * sum of squares; one block per lambda
__global__ void __launch_bounds__(LAM_THREADS) k_lambda_final(int nblk, double scale, const z64* __restrict__ part,
                                                              z64* __restrict__ out) {
  __shared__ z64 red[2 * LAM_THREADS];
  const int l = blockIdx.x;
  z64 s1 = mk(0.0, 0.0), s2 = mk(0.0, 0.0);
  for (int b = threadIdx.x; b < nblk; b += LAM_THREADS) {
    s1 = cadd(s1, part[((long)l * nblk + b) * 2]);
    s2 = cadd(s2, part[((long)l * nblk + b) * 2 + 1]);
  }
  block_reduce2(s1, s2, red);
  if (threadIdx.x == 0) {
    out[l * 2] = cscale(s1, scale);
    out[l * 2 + 1] = cscale(s2, -scale);
  }
}

void launch_lambda_sums(long npts, int nl, int nblk, double scale, const double* lams, const z64* chi, z64* part,
                        z64* out, hipStream_t s) {
  hipLaunchKernelGGL(k_lambda_partial, dim3(nblk, nl), dim3(LAM_THREADS), 0, s, npts, lams, chi, part);
  hipLaunchKernelGGL(k_lambda_final, dim3(nl), dim3(LAM_THREADS), 0, s, nblk, scale, part, out);
}

__global__ void k_lambda_apply(long npts, double lam, const z64* __restrict__ chi, z64* __restrict__ out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const z64 c = chi[p];
  out[p] = cdiv(c, mk(1.0 + lam * c.re, lam * c.im));
}

void launch_lambda_apply(long npts, double lam, const z64* chi, z64* out, hipStream_t s) {
  hipLaunchKernelGGL(k_lambda_apply, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, npts, lam, chi, out);
}

}  // namespace adga
