// The lambda correction of the final susceptibility (no counterpart in the reference): C-ABI entry points
// adga_lambda_*.  The physical susceptibility of every (q, omega) point stays on the device; each evaluation
// of the sum rule is one pair of FP64 reductions there (lambda.hip), and lambda is found on the host by a
// safeguarded Newton iteration on f(lambda) = Re S(lambda) - target, S(lambda) = 1/(beta nqp) sum chi^lambda:
// f decreases from +inf at the pole lambda_min = -1/max Re chi(q, omega = 0) and is convex there for
// positive susceptibilities, so a Newton step that leaves the bracket [a, b] (f(a) > 0 > f(b)) is replaced
// by bisection.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "adga_amd.h"
#include "adga_kernels.h"

using adga::z64;

struct adga_lambda {
  int n = 0, nqp = 0, nw = 0;
  long npts = 0;
  double beta = 1.0;
  double lmin = 0.0;       // the pole of the largest omega = 0 susceptibility
  z64* chi = nullptr;      // chi_phys of every point, omega-major (nqp per omega)
  z64* loc = nullptr;      // chi_phys of the local susceptibility per omega
  z64* part = nullptr;     // partial sums
  z64* out = nullptr;      // final sums
  double* lams = nullptr;  // the lambdas of one call
  int cap = 0;             // lambdas the buffers hold
  int nblk = 256;
  hipStream_t s = nullptr;
  std::string err;
};

namespace {

struct LamError {
  hipError_t e;
  const char* what;
};
struct NonFinite {};  // a sum of the susceptibilities is not finite

#define LCK(x)                                    \
  do {                                            \
    hipError_t e_ = (x);                          \
    if (e_ != hipSuccess) throw LamError{e_, #x}; \
  } while (0)

template <typename F>
int lam_guarded(adga_lambda* h, F&& f) {
  try {
    return f();
  } catch (const LamError& le) {
    if (h) h->err = std::string("HIP error in ") + le.what + ": " + hipGetErrorString(le.e);
    return le.e == hipErrorOutOfMemory ? ADGA_E_NOMEM : ADGA_E_HIP;
  } catch (const NonFinite&) {
    if (h) h->err = "the susceptibilities' sum is not finite";
    return ADGA_E_NONFINITE;
  } catch (const std::exception& ex) {
    if (h) h->err = std::string("host exception: ") + ex.what();
    return ADGA_E_HIP;
  }
}

template <typename T>
void lalloc(T** p, size_t count) {
  LCK(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
}
template <typename T>
void lfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

void release(adga_lambda* h) {
  lfree(h->chi);
  lfree(h->loc);
  lfree(h->part);
  lfree(h->out);
  lfree(h->lams);
  if (h->s) (void)hipStreamDestroy(h->s);
  h->s = nullptr;
}

void ensure_cap(adga_lambda* h, int nl) {
  if (nl <= h->cap) return;
  lfree(h->part);
  lfree(h->out);
  lfree(h->lams);
  lalloc(&h->part, (size_t)nl * h->nblk * 2);
  lalloc(&h->out, (size_t)nl * 2);
  lalloc(&h->lams, (size_t)nl);
  h->cap = nl;
}

// S and S' (complex, interleaved) of nl lambdas over `npts` values at `chi`, scaled by `scale`
void sums(adga_lambda* h, const z64* chi, long npts, double scale, int nl, const double* lams, z64* res) {
  ensure_cap(h, nl);
  LCK(hipMemcpyAsync(h->lams, lams, sizeof(double) * nl, hipMemcpyHostToDevice, h->s));
  adga::launch_lambda_sums(npts, nl, h->nblk, scale, h->lams, chi, h->part, h->out, h->s);
  LCK(hipGetLastError());
  LCK(hipMemcpyAsync(res, h->out, sizeof(z64) * 2 * nl, hipMemcpyDeviceToHost, h->s));
  LCK(hipStreamSynchronize(h->s));
}

}  // namespace

extern "C" {

int adga_lambda_create(int ndim, int nqp, int iwbmax, double beta, const double* chi_qw, const double* chi_loc,
                       adga_lambda** out) {
  if (!out || !chi_qw || !chi_loc || ndim < 1 || ndim > 8 || nqp < 1 || iwbmax < 0 || !(beta > 0))
    return ADGA_E_ARG;
  *out = nullptr;
  adga_lambda* h = new (std::nothrow) adga_lambda();
  if (!h) return ADGA_E_NOMEM;
  h->n = ndim;
  h->nqp = nqp;
  h->nw = 2 * iwbmax + 1;
  h->npts = (long)nqp * h->nw;
  h->beta = beta;
  const size_t n4 = (size_t)ndim * ndim * ndim * ndim;
  const int rc = lam_guarded(h, [&]() -> int {
    LCK(hipStreamCreate(&h->s));
    lalloc(&h->chi, (size_t)h->npts);
    lalloc(&h->loc, (size_t)h->nw);
    z64* full = nullptr;
    lalloc(&full, n4 * std::max<long>(h->npts, h->nw));
    try {
      LCK(hipMemcpyAsync(full, chi_qw, sizeof(z64) * n4 * h->npts, hipMemcpyHostToDevice, h->s));
      adga::launch_chi_phys(ndim, h->npts, full, h->chi, h->s);
      LCK(hipMemcpyAsync(full, chi_loc, sizeof(z64) * n4 * h->nw, hipMemcpyHostToDevice, h->s));
      adga::launch_chi_phys(ndim, h->nw, full, h->loc, h->s);
      LCK(hipGetLastError());
      // the pole: omega = 0 is the (iwbmax + 1)-th omega of the omega-major list
      std::vector<z64> c0(nqp);
      LCK(hipMemcpyAsync(c0.data(), h->chi + (long)iwbmax * nqp, sizeof(z64) * nqp, hipMemcpyDeviceToHost, h->s));
      LCK(hipStreamSynchronize(h->s));
      lfree(full);
      double cmax = -INFINITY;
      for (const z64& c : c0) cmax = std::max(cmax, c.re);
      if (!(cmax > 0)) {
        h->err = "the omega = 0 susceptibility has no positive value: no pole bounds lambda";
        return ADGA_E_ARG;
      }
      h->lmin = -1.0 / cmax;
    } catch (...) {
      lfree(full);
      throw;
    }
    return ADGA_OK;
  });
  if (rc != ADGA_OK) {
    release(h);
    delete h;
    return rc;
  }
  *out = h;
  return ADGA_OK;
}

int adga_lambda_sums(adga_lambda* h, int nl, const double* lambdas, double* s, double* ds) {
  if (!h || nl < 1 || !lambdas || !s || !ds) return ADGA_E_ARG;
  return lam_guarded(h, [&]() -> int {
    std::vector<z64> res((size_t)2 * nl);
    sums(h, h->chi, h->npts, 1.0 / (h->beta * h->nqp), nl, lambdas, res.data());
    for (int l = 0; l < nl; ++l) {
      s[2 * l] = res[2 * l].re;
      s[2 * l + 1] = res[2 * l].im;
      ds[2 * l] = res[2 * l + 1].re;
      ds[2 * l + 1] = res[2 * l + 1].im;
    }
    return ADGA_OK;
  });
}

int adga_lambda_target(adga_lambda* h, double* target) {
  if (!h || !target) return ADGA_E_ARG;
  return lam_guarded(h, [&]() -> int {
    z64 res[2];
    const double zero = 0.0;
    sums(h, h->loc, h->nw, 1.0 / h->beta, 1, &zero, res);
    *target = res[0].re;
    return ADGA_OK;
  });
}

int adga_lambda_solve(adga_lambda* h, double target, double tol, int maxit, double* lambda, int* iters) {
  if (!h || !lambda || !(tol > 0) || maxit < 1 || !std::isfinite(target)) return ADGA_E_ARG;
  return lam_guarded(h, [&]() -> int {
    const double scale = 1.0 / (h->beta * h->nqp);
    auto f = [&](double lam, double* df) {
      z64 res[2];
      sums(h, h->chi, h->npts, scale, 1, &lam, res);
      if (!std::isfinite(res[0].re) || !std::isfinite(res[1].re)) throw NonFinite{};
      *df = res[1].re;
      return res[0].re - target;
    };
    // bracket: a just above the pole (f -> +inf), b where f < 0
    double a = h->lmin, b = 1.0, df = 0.0;
    double fb = f(b, &df);
    int it = 0;
    while (fb > 0) {
      if (++it > 200) {
        h->err = "no lambda above the pole brings the sum below the target";
        return ADGA_E_NOCONV;
      }
      a = b;
      b = b > 0 ? 4.0 * b : 1.0;
      fb = f(b, &df);
    }
    double x = b, fx = fb, dfx = df;
    const double ftol = tol * std::max(1.0, std::fabs(target));
    for (int k = 0; k < maxit; ++k) {
      if (iters) *iters = k;
      if (std::fabs(fx) <= ftol) {
        *lambda = x;
        return ADGA_OK;
      }
      double xn = (dfx < 0) ? x - fx / dfx : 0.5 * (a + b);
      if (!(xn > a && xn < b)) xn = 0.5 * (a + b);  // Newton left the bracket: bisect
      x = xn;
      fx = f(x, &dfx);
      if (fx > 0)
        a = x;
      else
        b = x;
      if (b - a <= 4e-16 * std::max(1.0, std::fabs(x))) {
        if (iters) *iters = k + 1;
        *lambda = x;
        return ADGA_OK;
      }
    }
    if (iters) *iters = maxit;
    *lambda = x;
    h->err = "lambda: no convergence within maxit iterations";
    return ADGA_E_NOCONV;
  });
}

int adga_lambda_chi(adga_lambda* h, double lambda, double* chi_phys) {
  if (!h || !chi_phys) return ADGA_E_ARG;
  return lam_guarded(h, [&]() -> int {
    z64* tmp = nullptr;
    lalloc(&tmp, (size_t)h->npts);
    adga::launch_lambda_apply(h->npts, lambda, h->chi, tmp, h->s);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(chi_phys, tmp, sizeof(z64) * h->npts, hipMemcpyDeviceToHost, h->s);
    if (e == hipSuccess) e = hipStreamSynchronize(h->s);
    lfree(tmp);
    LCK(e);
    return ADGA_OK;
  });
}

int adga_lambda_last_error(adga_lambda* h, char* buf, int len) {
  if (!h || !buf || len < 1) return ADGA_E_ARG;
  strncpy(buf, h->err.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return ADGA_OK;
}

int adga_lambda_destroy(adga_lambda* h) {
  if (!h) return ADGA_OK;
  release(h);
  delete h;
  return ADGA_OK;
}

}  // extern "C"
