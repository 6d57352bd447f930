// Derivatives of the hot path along caller-given directions of its inputs (the tangent-linear form of
// bubble -> BSE -> chi / Sigma): the C-ABI entry points adga_set_tangents, adga_set_local_omega_tangent,
// adga_run_points_tangent (api.cpp) and adga_get_sigma_tangent, and the per-batch work they share.  The
// derivative of a batch reuses the batch's forward data (S, C, gamma, eta, m_tot and the complex128 LU
// factors): each direction costs two bubble passes (product rule: the bubble is formed again, not
// stored), a right-hand side and a solve with the existing factors, chi's contraction and two EOM passes.
#include <cstring>

#include "ctx.h"

using namespace adga_host;

namespace {

void ensure_tangent_batch(adga_ctx* c, int nb) {
  if (nb <= c->t_cap) return;
  for (z64** p : {&c->t_dchi0q, &c->t_dS, &c->t_Z1, &c->t_Z2, &c->t_B, &c->t_dY, &c->t_dgamq, &c->t_dmt,
                  &c->t_dchi, &c->t_dbub})
    dfree(*p);
  const Geom& g = c->g;  // the largest box
  const size_t per_pt = (size_t)g.nf * g.n4, pmat = (size_t)g.N * g.n2;
  dalloc(&c->t_dchi0q, nb * per_pt);
  dalloc(&c->t_dS, nb * per_pt);
  dalloc(&c->t_Z1, 2 * nb * pmat);
  dalloc(&c->t_Z2, 2 * nb * pmat);
  dalloc(&c->t_B, 2 * nb * pmat);
  dalloc(&c->t_dY, 2 * nb * pmat);
  dalloc(&c->t_dgamq, 2 * nb * pmat);
  dalloc(&c->t_dmt, nb * pmat);
  dalloc(&c->t_dchi, (size_t)c->ndir * 2 * nb * g.n4);
  dalloc(&c->t_dbub, (size_t)c->ndir * nb * g.n4);
  c->t_cap = nb;
}

void download(double* h, const z64* d, size_t count, hipStream_t s) {
  CK(hipMemcpyAsync(h, d, count * sizeof(z64), hipMemcpyDeviceToHost, s));
}

}  // namespace

namespace adga_host {

void free_tangent(adga_ctx* c) {
  for (z64** p : {&c->dgt, &c->dgiw, &c->du, &c->dut, &c->t_dw, &c->t_dwinv, &c->t_dcd, &c->t_dcm, &c->t_dgamd,
                  &c->t_dgamm, &c->t_dchi0q, &c->t_dS, &c->t_Z1, &c->t_Z2, &c->t_B, &c->t_dY, &c->t_dgamq,
                  &c->t_dmt, &c->t_dchi, &c->t_dbub, &c->t_sig, &c->t_macc, &c->t_dghat, &c->t_mwork,
                  &c->t_dbub_all, &c->t_bub_p})
    dfree(*p);
  dfree(c->dmu);
  c->t_bub_valid = false;
  c->ndir = 0;
  c->t_cap = 0;
  c->t_omega = false;
}

void tangent_batch(adga_ctx* c, int nb, int p0, int npts, const std::vector<int>& keep, int nfailed,
                   const TangentOut& out) {
  const Geom& g = c->gw;
  const Geom& G = c->g;
  hipStream_t s = c->stream;
  ensure_tangent_batch(c, nb);
  const int nmat = 2 * nb;
  const size_t tab = (size_t)G.nw * G.nkp * G.n2, loc = (size_t)G.n2 * G.nf, cc = (size_t)G.N * G.N,
               pm = (size_t)G.n2 * G.N, nsig = (size_t)G.n2 * G.nf * G.nkp_eom;
  for (int dir = 0; dir < c->ndir; ++dir) {
    const z64* dgt = c->dgt + dir * tab;
    const z64 *dw = c->t_dw + dir * loc, *dwinv = c->t_dwinv + dir * loc;
    const z64 *dcd = c->t_dcd + dir * cc, *dcm = c->t_dcm + dir * cc;
    const z64 *dgamd = c->t_dgamd + dir * pm, *dgamm = c->t_dgamm + dir * pm;
    // dchi0q = bubble(dG(k, nu), G(k - q, nu - omega)) + bubble(G(k, nu), dG(k - q, nu - omega)): for every q
    // of the omega at once by k-grid convolution (formed again, not stored from the forward pass, and kept
    // for the omega's next batches), or as direct sums per batch, as the forward bubble's mode chooses
    const bool t_hit = c->t_bub_valid && c->t_bub_iwb == c->iwb && c->t_bub_iwf == g.iwf;
    if (c->t_dghat && (c->bubble_mode == 1 || (c->bubble_mode == 2 && (t_hit || nb >= 64)))) {
      const size_t nall = (size_t)G.nf * G.n4 * G.nkp;  // sized for the largest box
      if (!t_hit) {
        if (dir == 0) {
          if (!c->t_dbub_all) {
            dalloc(&c->t_dbub_all, nall * c->ndir);
            dalloc(&c->t_bub_p, nall);
          }
          for (int d = 0; d < c->ndir; ++d)
            adga::launch_bubble_all_tangent(g, c->iwb, c->tw, c->ghat, c->t_dghat + d * tab, c->t_bub_p,
                                            c->t_dbub_all + d * nall, s);
          c->t_bub_valid = true;
          c->t_bub_iwb = c->iwb;
          c->t_bub_iwf = g.iwf;
        }
      }
      adga::launch_bubble_gather(g, nb, c->qidx, c->t_dbub_all + dir * nall, c->t_dchi0q, s);
    } else {
      adga::launch_bubble(g, c->iwb, nb, c->qidx, dgt, c->t_dchi0q, s, c->gt, false);
      adga::launch_bubble(g, c->iwb, nb, c->qidx, c->gt, c->t_dchi0q, s, dgt, true);
    }
    adga::launch_dsmat(g, nb, c->chi0q, c->t_dchi0q, c->winv, dwinv, c->t_dS, s);
    // d eta: M d eta^T = C^T Z1 + dC^T Z2, solved with the complex128 factors of the forward solve
    adga::launch_tangent_rhs(g, nb, c->S, c->t_dS, c->cd, c->cm, dcd, dcm, c->opgd, c->opgm, dgamd, dgamm, c->Y,
                             c->t_Z1, c->t_Z2, c->t_B, s);
    if (!adga::launch_lu_solve<double>(nmat, g.N, g.n2, c->ldm, (const z64*)c->M, c->piv, (const z64*)c->inv, c->t_B,
                                       c->t_dY, false, s))
      throw HipError{hipErrorInvalidValue, "tangent solve: BSE size outside the solve kernel's range"};
    adga::launch_dchi(g, nb, c->chi0q, c->t_dchi0q, c->chi0w, dw, c->gamd, c->gamm, dgamd, dgamm, c->Y, c->t_dY,
                      c->t_dchi + (size_t)dir * 2 * nb * g.n4, c->t_dbub + (size_t)dir * nb * g.n4, s);
    if (!c->do_eom) continue;
    // d m_tot = m_tot(U, U~; d eta, d gamma^q) + m_tot(dU, dU~; eta, gamma^q), with
    // d gamma^q_d = sum_nu (dS C_d + S dC_d)[nu rows]
    adga::launch_gamq_z(g, nb, c->t_dS, c->t_Z1, s);
    adga::resid_gemm(g, nb, 0, 1, c->cd, c->cd, c->t_Z1, c->t_dgamq, nullptr, nullptr, s);
    adga::launch_gamq_z(g, nb, c->S, c->t_Z1, s);
    adga::resid_gemm(g, nb, 0, 1, dcd, dcd, c->t_Z1, c->t_dgamq, nullptr, nullptr, s, true);
    adga::launch_mtot(g, nb, c->u, c->ut, c->t_dY, c->t_dgamq, c->t_dmt, s);  // d eta = dY
    adga::launch_mtot(g, nb, c->du + (size_t)dir * G.n4, c->dut + (size_t)dir * G.n4, c->Y, c->gamq, c->t_dmt, s,
                      nullptr, nullptr, nullptr, true);
    if (nfailed > 0) adga::launch_zero_points(nb, (long)g.n2 * g.N, c->keep, c->t_dmt, s);  // keep flags on the device
    // d Sigma_nl = sum_q [d m_tot . G(k - q) + m_tot . dG(k - q)]: with the EOM as k-grid convolutions,
    // d m_tot is accumulated on the q grid beside m_tot and both terms are formed when the omega's EOM is
    // flushed; else as direct sums per batch
    if (c->eom_mode == 1 && c->t_macc) {
      adga::launch_eom_scatter(g, nb, c->qidx, c->t_dmt, c->t_macc + dir * (size_t)G.iwf * G.n4 * G.nkp, s,
                               c->q_distinct);
    } else {
      z64* sig = c->t_sig + dir * nsig;
      adga::launch_eom_direct(g, c->iwb, nb, c->qidx, c->keom, c->t_dmt, c->gt, sig, s);
      adga::launch_eom_direct(g, c->iwb, nb, c->qidx, c->keom, c->mt, dgt, sig, s);
    }
  }
  CK(hipGetLastError());
  // outputs (ndim^2, ndim^2, npts, ndir)
  std::vector<z64> h((size_t)c->ndir * 2 * nb * g.n4), hb((size_t)c->ndir * nb * g.n4);
  CK(hipMemcpyAsync(h.data(), c->t_dchi, sizeof(z64) * h.size(), hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(hb.data(), c->t_dbub, sizeof(z64) * hb.size(), hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  (void)keep;
  for (int dir = 0; dir < c->ndir; ++dir)
    for (int b = 0; b < nb; ++b) {
      const size_t o = ((size_t)dir * npts + p0 + b) * 2 * g.n4;  // interleaved doubles
      const z64* src = &h[((size_t)dir * 2 * nb + 2 * b) * g.n4];
      if (out.dchid) memcpy(out.dchid + o, src, sizeof(z64) * g.n4);
      if (out.dchim) memcpy(out.dchim + o, src + g.n4, sizeof(z64) * g.n4);
      if (out.dbub) memcpy(out.dbub + o, &hb[((size_t)dir * nb + b) * g.n4], sizeof(z64) * g.n4);
    }
}

}  // namespace adga_host

extern "C" {

int adga_set_tangents(adga_ctx* c, int ndir, const double* dhk, const double* dsiw, const double* dgiw,
                      const double* dmu, const double* du, const double* dut) {
  if (!c || ndir < 1 || ndir > 16 || !dhk || !dsiw || !dgiw || !dmu || !du || !dut) return ADGA_E_ARG;
  if (c->eom_pending) return fail(c, ADGA_E_STATE, "tangents cannot change while an omega is pending (adga_get_sigma)");
  if (c->bubble_range != 0) return fail(c, ADGA_E_STATE, "the derivatives are formed with the bubble on the box");
  if (c->g.n > 5) return fail(c, ADGA_E_ARG, "derivatives: orbital count above 5");
  free_tangent(c);  // ndir = 0 until every buffer below is in place
  const int rc = guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    const size_t nh = (size_t)g.n2 * g.nkp * ndir, ns = (size_t)2 * g.iwmax * g.n * ndir;
    DevBuf<> dh(nh), dsw(ns);
    upload(dh.p, dhk, nh, s);
    upload(dsw.p, dsiw, ns, s);
    dalloc(&c->dgiw, ns);
    upload(c->dgiw, dgiw, ns, s);
    dalloc(&c->dmu, (size_t)ndir);
    CK(hipMemcpyAsync(c->dmu, dmu, sizeof(double) * ndir, hipMemcpyHostToDevice, s));
    dalloc(&c->du, (size_t)g.n4 * ndir);
    dalloc(&c->dut, (size_t)g.n4 * ndir);
    upload(c->du, du, (size_t)g.n4 * ndir, s);
    upload(c->dut, dut, (size_t)g.n4 * ndir, s);
    dalloc(&c->dgt, (size_t)g.nw * g.nkp * g.n2 * ndir);
    if (!adga::launch_dgtable(g, ndir, c->gt, dh.p, dsw.p, c->dmu, c->dgt, s))
      return fail(c, ADGA_E_ARG, "derivatives: orbital count above 5");
    c->ndir = ndir;  // the per-direction buffers below are sized by it
    const size_t loc = (size_t)g.n2 * g.nf * ndir, cc = (size_t)g.N * g.N * ndir, pm = (size_t)g.n2 * g.N * ndir;
    dalloc(&c->t_dw, loc);
    dalloc(&c->t_dwinv, loc);
    dalloc(&c->t_dcd, cc);
    dalloc(&c->t_dcm, cc);
    dalloc(&c->t_dgamd, pm);
    dalloc(&c->t_dgamm, pm);
    const size_t nsig = (size_t)g.n2 * g.nf * g.nkp_eom * ndir;
    dalloc(&c->t_sig, nsig);
    CK(hipMemsetAsync(c->t_sig, 0, sizeof(z64) * nsig, s));
    const bool eom_dft = c->do_eom && c->eom_mode == 1;
    if (eom_dft || c->bubble_mode != 0) {
      // the derivatives by k-grid convolution (EOM, bubble): each direction's dG table transformed once
      conv_tables(c);
      const size_t ngt = (size_t)g.nw * g.n2 * g.nkp;
      dalloc(&c->t_dghat, ngt * ndir);
      DevBuf<> tmp(ngt);
      for (int d = 0; d < ndir; ++d) adga::launch_ghat(g, c->tw, c->dgt + d * ngt, tmp.p, c->t_dghat + d * ngt, s);
    }
    if (eom_dft) {
      const size_t nm = (size_t)g.iwf * g.n4 * g.nkp;
      dalloc(&c->t_macc, nm * ndir);
      CK(hipMemsetAsync(c->t_macc, 0, sizeof(z64) * nm * ndir, s));
      dalloc(&c->t_mwork, nm);
    }
    CK(hipGetLastError());
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
  if (rc != ADGA_OK) free_tangent(c);  // no half-set directions
  return rc;
}

int adga_set_local_omega_tangent(adga_ctx* c, const double* dxd, const double* dxm, double* dchid_out,
                                 double* dchim_out, double* dbubloc_out) {
  if (!c) return ADGA_E_ARG;
  if (c->ndir == 0) return fail(c, ADGA_E_STATE, "adga_set_tangents must be called first");
  if (!c->have_omega) return fail(c, ADGA_E_STATE, "adga_set_local_omega must be called first");
  return guarded(c, [&]() -> int {
    const Geom& g = c->gw;
    const Geom& G = c->g;
    hipStream_t s = c->stream;
    const size_t nn = (size_t)g.N * g.N;
    const size_t loc = (size_t)G.n2 * G.nf, cc = (size_t)G.N * G.N, pm = (size_t)G.n2 * G.N;
    const size_t ng = (size_t)2 * G.iwmax * G.n;
    DevBuf<> dx((dxd || dxm) ? 2 * nn : 1), out((size_t)3 * G.n4 * c->ndir);
    for (int dir = 0; dir < c->ndir; ++dir) {
      if (dxd) upload(dx.p, dxd + 2 * nn * dir, nn, s);
      if (dxm) upload(dx.p + nn, dxm + 2 * nn * dir, nn, s);
      z64* o = out.p + (size_t)3 * G.n4 * dir;
      adga::launch_local_tangent(g, c->iwb, c->giw, c->dgiw + dir * ng, c->xd, c->xm, dxd ? dx.p : nullptr,
                                 dxm ? dx.p + nn : nullptr, c->winv, c->gamd, c->gamm, c->chi0w, c->t_dw + dir * loc,
                                 c->t_dwinv + dir * loc, c->t_dcd + dir * cc, c->t_dcm + dir * cc,
                                 c->t_dgamd + dir * pm, c->t_dgamm + dir * pm, o, o + G.n4, o + 2 * G.n4, s);
      CK(hipGetLastError());
      CK(hipStreamSynchronize(s));  // dx is reused by the next direction
    }
    std::vector<z64> h((size_t)3 * G.n4 * c->ndir);
    CK(hipMemcpyAsync(h.data(), out.p, sizeof(z64) * h.size(), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (int dir = 0; dir < c->ndir; ++dir) {
      const z64* src = &h[(size_t)3 * G.n4 * dir];
      const size_t o = (size_t)2 * G.n4 * dir;
      if (dchid_out) memcpy(dchid_out + o, src, sizeof(z64) * G.n4);
      if (dchim_out) memcpy(dchim_out + o, src + G.n4, sizeof(z64) * G.n4);
      if (dbubloc_out) memcpy(dbubloc_out + o, src + 2 * G.n4, sizeof(z64) * G.n4);
    }
    c->t_omega = true;
    c->t_iwb = c->iwb;
    c->t_iwf = g.iwf;
    return ADGA_OK;
  });
}

int adga_get_sigma_tangent(adga_ctx* c, double* dsigma_nl) {
  if (!c || !dsigma_nl) return ADGA_E_ARG;
  if (c->ndir == 0) return fail(c, ADGA_E_STATE, "adga_set_tangents must be called first");
  return guarded(c, [&]() -> int {
    const Geom& g = c->g;
    hipStream_t s = c->stream;
    const size_t nsig = (size_t)g.n2 * g.nf * g.nkp_eom;
    flush_eom(c);  // the pending omega's EOM and its derivatives
    for (int dir = 0; dir < c->ndir; ++dir) adga::launch_fill_neg(g, g.nkp_eom, c->t_sig + dir * nsig, s);
    CK(hipGetLastError());
    download(dsigma_nl, c->t_sig, nsig * c->ndir, s);
    CK(hipStreamSynchronize(s));
    return ADGA_OK;
  });
}

}  // extern "C"
