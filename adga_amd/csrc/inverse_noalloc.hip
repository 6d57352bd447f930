// adga_inverse_matrix_z_noalloc: the inverse of adga_inverse_matrix_z (same operation list, same
// kernels, same layout) with every device buffer it needs carved from one file-scope device array, so
// that a call makes no device allocation.  Each GPU the library runs on holds its own copy of the array
// (about 34.6 MB at the largest n) until the process ends; every call on a GPU shares it, so calls that
// overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "adga_amd.h"
#include "adga_kernels.h"

namespace {

using adga::z64;

constexpr int NA_MAX_N = ADGA_INVERSE_NOALLOC_MAX_N;
constexpr long NA_LDM = (2L * NA_MAX_N + 15) / 16 * 16;        // row stride: the matrix and its n right-hand sides
constexpr long NA_M = (long)NA_MAX_N * NA_LDM;                  // complex elements of the factored block
constexpr long NA_INV = (NA_MAX_N + 31) / 32 * 2L * 32 * 32;    // panel inverses (lu_inv_elems at NA_MAX_N)
constexpr long NA_INTS = NA_MAX_N + 1;                          // row permutation, pivot status
constexpr long NA_ELEMS = NA_M + NA_INV + (NA_INTS * (long)sizeof(int) + sizeof(z64) - 1) / sizeof(z64);
constexpr int NA_OPS_CAP = 4096;                                // ints of the operation list (256 operations at n = 1024)
constexpr int NA_MAX_DEV = 64;

}  // namespace

static __device__ __align__(256) z64 scratch[NA_ELEMS];

namespace {

std::mutex na_mu;
void* na_ptr[NA_MAX_DEV] = {};

// Device address of `scratch` on the current device, looked up once per device.
hipError_t scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= NA_MAX_DEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(na_mu);
  if (!na_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    na_ptr[dev] = p;
  }
  *out = na_ptr[dev];
  return hipSuccess;
}

// columns [n, ldm) of the n rows: the identity right-hand sides (column n + r of row r), then padding
__global__ void k_identity_rhs(int n, int ldm, z64* __restrict__ M) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = ldm - n;
  if (idx >= (long)n * w) return;
  const int r = (int)(idx / w), c = (int)(idx % w);
  M[(long)r * ldm + n + c] = adga::mk(c == r ? 1.0 : 0.0, 0.0);
}

}  // namespace

extern "C" int adga_inverse_matrix_z_noalloc(int n, double* m, int* ierr) {
  if (n < 1 || !m || !ierr) return ADGA_E_ARG;
  *ierr = 0;
  if (n > NA_MAX_N) return ADGA_E_ARG;
  int ops[NA_OPS_CAP];
  const int W = 2 * n, ldm = (W + 15) / 16 * 16;
  const size_t nops = adga::lu_ops_factor_into(n, W, ops, NA_OPS_CAP);
  if (nops > (size_t)NA_OPS_CAP || adga::lu_inv_elems(n) > (size_t)NA_INV) return ADGA_E_ARG;
  void* base = nullptr;
  if (scratch_address(&base) != hipSuccess) return ADGA_E_HIP;
  z64* M = static_cast<z64*>(base);
  z64* inv = M + NA_M;
  int* piv = reinterpret_cast<int*>(inv + NA_INV);
  int* info = piv + NA_MAX_N;
  // row r of the row-major block = column r of the Fortran matrix: the factored matrix is m^T, and
  // solving m^T Y = 1 gives Y = m^-T, whose rows are the columns of m^-1
  if (hipMemcpy2D(M, sizeof(z64) * ldm, m, sizeof(z64) * n, sizeof(z64) * n, n, hipMemcpyHostToDevice) != hipSuccess)
    return ADGA_E_HIP;
  const long tot = (long)n * (ldm - n);
  hipLaunchKernelGGL(k_identity_rhs, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, nullptr, n, ldm, M);
  if (!adga::launch_lu<double>(1, n, W, ldm, ops, nops, M, piv, inv, info, nullptr)) return ADGA_E_ARG;
  if (hipGetLastError() != hipSuccess) return ADGA_E_HIP;
  if (hipMemcpy(ierr, info, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return ADGA_E_HIP;
  if (*ierr != 0) return ADGA_E_SINGULAR;
  if (hipMemcpy2D(m, sizeof(z64) * n, M + n, sizeof(z64) * ldm, sizeof(z64) * n, n, hipMemcpyDeviceToHost) !=
      hipSuccess)
    return ADGA_E_HIP;
  return ADGA_OK;
}
