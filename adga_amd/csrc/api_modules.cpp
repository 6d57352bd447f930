// Module-level entry points of the C ABI: one call per call of a reference module procedure, so that
// replacement Fortran modules with the reference's names and interfaces (adga_amd/fortran/dropin/) can
// forward to the GPU with the calling program unchanged.  They run on the context's stream and
// device tables; caller arrays are host memory in the reference's layouts.
#include <cstring>

#include "ctx.h"

using namespace adga_host;

namespace {

// device staging of a caller array of `count` complex values: the context keeps one buffer per slot
z64* stage(adga_ctx* c, int slot, size_t count) {
  if (c->stage_n[slot] < count) {
    dfree(c->stage[slot]);
    dalloc(&c->stage[slot], count);
    c->stage_n[slot] = count;
  }
  return c->stage[slot];
}

void download(double* h, const z64* d, size_t count, hipStream_t s) {
  CK(hipMemcpyAsync(h, d, count * sizeof(z64), hipMemcpyDeviceToHost, s));
}

bool iwb_ok(const adga_ctx* c, int iwb) { return iwb >= -c->g.iwbmax && iwb <= c->g.iwbmax; }

}  // namespace

extern "C" {

int adga_bubble_q(adga_ctx* c, int iwb, int q_k, double* chi0q) {
  if (!c || !chi0q) return ADGA_E_ARG;
  if (!iwb_ok(c, iwb)) return fail(c, ADGA_E_ARG, "iwb outside [-iwbmax_small, iwbmax_small]");
  if (q_k < 1 || q_k > c->g.nkp) return fail(c, ADGA_E_ARG, "q_k outside [1, nkp]");
  return guarded(c, [&]() -> int {
    const Geom g = bubble_geom(c);
    hipStream_t s = c->stream;
    const size_t ne = (size_t)g.nf * g.n4;
    int* qd = reinterpret_cast<int*>(stage(c, 0, 1));
    const int q0 = q_k - 1;
    CK(hipMemcpyAsync(qd, &q0, sizeof(int), hipMemcpyHostToDevice, s));
    z64* out = stage(c, 1, ne);
    adga::launch_bubble(g, iwb, 1, qd, c->gt, out, s);
    CK(hipGetLastError());
    download(chi0q, out, ne, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_chi0_loc(adga_ctx* c, int iwb, double* chi0w_diag, double* chi0w_inv_diag) {
  if (!c || (!chi0w_diag && !chi0w_inv_diag)) return ADGA_E_ARG;
  if (!iwb_ok(c, iwb)) return fail(c, ADGA_E_ARG, "iwb outside [-iwbmax_small, iwbmax_small]");
  return guarded(c, [&]() -> int {
    const Geom g = bubble_geom(c);
    hipStream_t s = c->stream;
    const size_t ne = (size_t)g.n2 * g.nf;
    z64* w = stage(c, 0, ne);
    z64* wi = stage(c, 1, ne);
    adga::launch_chi0_loc_diag(g, iwb, c->giw, w, wi, s);
    CK(hipGetLastError());
    if (chi0w_diag) download(chi0w_diag, w, ne, s);
    if (chi0w_inv_diag) download(chi0w_inv_diag, wi, ne, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_chi_qw(adga_ctx* c, const double* interm3, const double* chi0_sum, double* chi_qw) {
  if (!c || !interm3 || !chi0_sum || !chi_qw) return ADGA_E_ARG;
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    const size_t nx = (size_t)g.n2 * g.N, ny = (size_t)g.n4 * g.nf;
    z64* X = stage(c, 0, nx);
    z64* Y = stage(c, 1, ny);
    z64* chi = stage(c, 2, g.n4);
    upload(X, interm3, nx, s);
    upload(Y, chi0_sum, ny, s);
    upload(chi, chi_qw, g.n4, s);
    adga::launch_chi_qw(g, X, Y, chi, s);
    CK(hipGetLastError());
    download(chi_qw, chi, g.n4, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_eom_dynamic(adga_ctx* c, int iwb, int iq, const double* etaqd, const double* etaqm, const double* gammawd,
                     const double* gammaqd, const double* v, double* sigma_nl) {
  if (!c || !etaqd || !etaqm || !gammawd || !gammaqd || !sigma_nl) return ADGA_E_ARG;
  if (!iwb_ok(c, iwb)) return fail(c, ADGA_E_ARG, "iwb outside [-iwbmax_small, iwbmax_small]");
  if (iq < 1 || iq > c->g.nqp) return fail(c, ADGA_E_ARG, "iq outside [1, nqp]");
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    const size_t per = (size_t)g.n2 * g.N, nsig = (size_t)g.n2 * g.nf * g.nkp_eom;
    z64* eta = stage(c, 0, 2 * per);  // density then magnetic, as k_mtot reads them
    z64* gq = stage(c, 1, per);
    z64* gw = stage(c, 2, per);
    z64* mt = stage(c, 3, per);
    z64* sig = stage(c, 4, nsig);
    z64* vb = stage(c, 5, (size_t)g.n4 + 1);
    int* ints = reinterpret_cast<int*>(vb + g.n4);  // q position 0 and the q's k index
    const int hi[2] = {0, c->q0[iq - 1]};
    upload(eta, etaqd, per, s);
    upload(eta + per, etaqm, per, s);
    upload(gq, gammaqd, per, s);
    upload(gw, gammawd, per, s);
    if (v) upload(vb, v, g.n4, s);
    CK(hipMemcpyAsync(ints, hi, sizeof hi, hipMemcpyHostToDevice, s));
    upload(sig, sigma_nl, nsig, s);
    // k_mtot reads gamma^q_d of point b at slot 2b of a two-slot layout: one point, slot 0
    adga::launch_mtot(g, 1, c->u, c->ut, eta, gq, mt, s, ints, v ? vb : nullptr, gw);
    adga::launch_eom_direct(g, iwb, 1, ints + 1, c->keom, mt, c->gt, sig, s);
    adga::launch_fill_neg(g, g.nkp_eom, sig, s);
    CK(hipGetLastError());
    download(sigma_nl, sig, nsig, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_eom_dmft(adga_ctx* c, int iwb, const double* gammawd, double* sigma_dmft) {
  if (!c || !gammawd || !sigma_dmft) return ADGA_E_ARG;
  if (!iwb_ok(c, iwb)) return fail(c, ADGA_E_ARG, "iwb outside [-iwbmax_small, iwbmax_small]");
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    const size_t per = (size_t)g.n2 * g.N, nsd = (size_t)g.n2 * g.nf;
    z64* gw = stage(c, 0, per);
    z64* sd = stage(c, 1, nsd);
    upload(gw, gammawd, per, s);
    upload(sd, sigma_dmft, nsd, s);
    adga::launch_eom_local(g, iwb, false, c->u, gw, c->giw, c->ndmft, sd, s);
    CK(hipGetLastError());
    download(sigma_dmft, sd, nsd, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_eom_static(adga_ctx* c, int iq, const double* v, int nonlocal, double* sigma_dmft, double* sigma_hf) {
  if (!c || !sigma_dmft || !sigma_hf) return ADGA_E_ARG;
  if (iq < 1 || iq > c->g.nqp) return fail(c, ADGA_E_ARG, "iq outside [1, nqp]");
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    const size_t nsd = (size_t)g.n2 * g.nf, nhf = (size_t)g.n2 * g.nkp_eom;
    const bool vterm = nonlocal && v;
    z64* sd = stage(c, 0, nsd);
    z64* hf = stage(c, 1, nhf);
    z64* vb = stage(c, 2, (size_t)g.n4 + 1);
    int* ints = reinterpret_cast<int*>(vb + g.n4);
    if (iq == 1) {
      upload(sd, sigma_dmft, nsd, s);
      adga::launch_eom_static_local(g, c->u, c->ndmft, sd, s);
      download(sigma_dmft, sd, nsd, s);
    }
    if (vterm) {
      const int hi[2] = {0, c->q0[iq - 1]};
      upload(vb, v, g.n4, s);
      CK(hipMemcpyAsync(ints, hi, sizeof hi, hipMemcpyHostToDevice, s));
      upload(hf, sigma_hf, nhf, s);
      ensure_nfock(c);
      if (iq == 1) adga::launch_eom_hartree_v(g, vb, c->ndmft, hf, s);
      adga::launch_eom_fock(g, 1, ints, ints + 1, c->keom, vb, c->nfock, hf, s);
      download(sigma_hf, hf, nhf, s);
    }
    CK(hipGetLastError());
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_get_gkiw(adga_ctx* c, int ik, int iw, const double* skiw, double* gkiw) {
  if (!c || !gkiw) return ADGA_E_ARG;
  if (ik < 1 || ik > c->g.nkp) return fail(c, ADGA_E_ARG, "ik outside [1, nkp]");
  if (!skiw && (iw < -c->g.iwmax || iw >= c->g.iwmax)) return fail(c, ADGA_E_ARG, "iw outside [-iwmax, iwmax)");
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    z64* out = stage(c, 0, g.n2);
    z64* sk = nullptr;
    if (skiw) {
      sk = stage(c, 1, g.n2);
      upload(sk, skiw, g.n2, s);
    }
    const int rc = op_finish(c, adga::launch_gkiw(g, ik - 1, iw, c->hk, c->siw, c->dc, sk, out, s), "get_gkiw");
    if (rc != ADGA_OK) return rc;
    download(gkiw, out, g.n2, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_get_ndmft(adga_ctx* c, double* n_dmft) {
  if (!c || !n_dmft) return ADGA_E_ARG;
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    adga::launch_ndmft(g, c->giw, c->ndmft, s);  // the context's n_dmft (the local Hartree term's input) too
    CK(hipGetLastError());
    download(n_dmft, c->ndmft, g.n, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

int adga_add_siw_dmft(adga_ctx* c, double* sigma_sum) {
  if (!c || !sigma_sum) return ADGA_E_ARG;
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    const size_t ns = (size_t)g.n2 * g.nf * g.nkp_eom;
    z64* ss = stage(c, 0, ns);
    upload(ss, sigma_sum, ns, s);
    adga::launch_add_siw(g, g.nkp_eom, c->siw, ss, s);
    CK(hipGetLastError());
    download(sigma_sum, ss, ns, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

}  // extern "C"
