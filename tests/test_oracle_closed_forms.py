"""Pins of the oracle that do not depend on the reference binary: closed forms,
limits and invariants chosen so that a dropped term, a wrong sign, a wrong
frequency shift or a transposed operand fails at least one of them."""
import dataclasses

import numpy as np
import pytest

from conftest import rel_err
from adga_amd import synth
import adga_oracle


def _inp(name, **kw):
    return synth.build_inputs(synth.workload(name, **kw))


def test_adjugate_inverse_matches_lapack():
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 4, 5):
        a = rng.normal(size=(50, n, n)) + 1j * rng.normal(size=(50, n, n)) + 3 * np.eye(n)
        assert rel_err(adga_oracle.inverse_small(a), np.linalg.inv(a)) < 1e-13


def test_k_minus_q_matches_vector_arithmetic():
    nk = (4, 3, 5)
    q = np.array([1, 7, 23, 60])
    t = adga_oracle.kmq_table(nk, q)
    for ik in range(np.prod(nk)):
        kv = np.array([ik // (nk[1] * nk[2]), (ik // nk[2]) % nk[1], ik % nk[2]])
        for j, qq in enumerate(q - 1):
            qv = np.array([qq // (nk[1] * nk[2]), (qq // nk[2]) % nk[1], qq % nk[2]])
            d = (kv - qv) % np.array(nk)
            assert t[ik, j] == d[2] + d[1] * nk[2] + d[0] * nk[1] * nk[2]
    # and the 1-based reference form
    assert synth.k_minus_q(1, 1, nk) == 1


def test_g_table_matsubara_symmetry():
    """G(k, -nu-1) = G(k, nu)^H for Hermitian H(k) and Sigma(-nu-1) = Sigma(nu)^*."""
    inp = _inp("t2g_k4")
    hp = adga_oracle.HotPath(inp)
    for w in range(0, hp.ghi + 1):
        if -w - 1 < hp.glo:
            continue
        assert rel_err(hp.G(-w - 1), np.conj(np.transpose(hp.G(w), (0, 2, 1)))) < 1e-14


def test_F_zero_limit():
    """chi_loc = chi0^w exactly: chi0^w F = 0, gamma = 0, A = 1, eta = 0, so the non-local chi parts and
    Sigma_nl vanish and the local chi is zero."""
    inp = _inp("t2g_k4", noise=0.0)
    hp = adga_oracle.HotPath(inp)
    pts = [(w, q) for w in (-1, 0, 2) for q in range(0, 64, 9)]
    out = hp.run(pts, lambda w, c: synth.chi_loc(inp, w, c))
    scale = np.abs(out["bubble_nl"]).max()
    assert scale > 1e-6
    for k in ("chi_d", "chi_m"):
        assert np.abs(out[k]).max() < 1e-12 * max(1.0, scale)
    assert np.abs(out["sigma_nl"]).max() < 1e-14
    for w in out["chi_loc_d"]:
        assert np.abs(out["chi_loc_d"][w]).max() < 1e-14


def test_local_limit_one_k_point():
    """One k point and orbital-diagonal H: G(k) is the local g, chi0^q = chi0^w, S = 0, so A = 1,
    eta = 0, chi_nl = 0 and bubble_nl = 0 whatever the vertex.  A wrong frequency shift in
    G(k-q, nu-omega) or a transposed chi0 breaks it."""
    wl = synth.workload("two_k4", nk=(1, 1, 1))
    inp = synth.build_inputs(wl)
    inp.hk[0, 1, 0] = inp.hk[1, 0, 0] = 0.0
    # rebuild g from the diagonal H so that giw is exactly G
    hp = adga_oracle.HotPath(inp)
    for iw in range(-inp.iwmax, inp.iwmax):
        g = hp.g_matrix(np.array([0]), iw)[0]
        inp.giw[iw + inp.iwmax] = np.diag(g)
    hp = adga_oracle.HotPath(inp)
    pts = [(w, 0) for w in range(-wl.iwbmax_small, wl.iwbmax_small + 1)]
    out = hp.run(pts, lambda w, c: synth.chi_loc(inp, w, c), keep=set(pts))
    for p in pts:
        r = out["points"][p]
        assert np.abs(r["eta_d"]).max() < 1e-12
        assert np.abs(r["eta_m"]).max() < 1e-12
    assert np.abs(out["chi_d"]).max() < 1e-12
    assert np.abs(out["bubble_nl"]).max() < 1e-14


def test_rank_one_vertex_sherman_morrison():
    """1 orbital, chi_loc(nu, nu') = chi0^w(nu) delta + c^2 chi0^w(nu) chi0^w(nu'):
    chi0^w F = u v^T with u = c chi0^w, v = c, so A = 1 - (S u) v^T and A^-1 follows from
    Sherman-Morrison.  With sigma = c sum_nu S_nu chi0^w(nu) and W = sum chi0^w:
      gamma^w = c^2 W, eta(nu') = (1 + c^2 W) c sigma / (1 - c sigma),
      chi_nl = beta^-2 [c^2 W sum (chi0^q - chi0^w) + eta sum chi0^q],  gamma^q_d = c sigma."""
    wl = synth.workload("sc1_k4")
    inp = synth.build_inputs(wl)
    hp = adga_oracle.HotPath(inp)
    c = 0.37
    for iwb, iq in ((0, 5), (1, 17), (-2, 42)):
        w = np.array([hp.chi0_loc(d - hp.iwf, iwb)[0] for d in range(hp.nf)])
        chil = np.diag(w) + c * c * np.outer(w, w)
        loc = hp.local_omega(iwb, chil, chil)
        r = hp.point(iq, iwb, loc)
        cq = r["chi0q"][0, 0, :]
        S = cq / w - 1.0
        sigma = c * np.sum(S * w)
        W = w.sum()
        eta = (1 + c * c * W) * c * sigma / (1 - c * sigma)
        assert rel_err(r["eta_d"][0], np.full(hp.nf, eta)) < 1e-12
        assert rel_err(r["gammaqd"][0], np.full(hp.nf, c * sigma)) < 1e-12
        chi = (c * c * W * np.sum(cq - w) + eta * np.sum(cq)) / inp.beta ** 2
        assert abs(r["chi_d"][0, 0] - chi) <= 1e-12 * abs(chi)


def test_bubble_brute_force_tiny():
    """Explicit loops over k with k - q from vector arithmetic and numpy inverses, 2 orbitals, 2^3 k."""
    wl = synth.workload("two_k4", nk=(2, 2, 2), iwfmax_small=2, iwbmax_small=1)
    inp = synth.build_inputs(wl)
    hp = adga_oracle.HotPath(inp)
    n = wl.n
    nu = inp.iw_data
    off = inp.iwmax
    for iq in range(8):
        for iwb in (-1, 0, 1):
            cq = hp.chi0q(iq, iwb)
            for d in range(wl.nf):
                iwf = d - wl.iwfmax_small
                acc = np.zeros((n, n, n, n), dtype=complex)
                qv = np.array([iq // 4, (iq // 2) % 2, iq % 2])
                for ik in range(8):
                    kv = np.array([ik // 4, (ik // 2) % 2, ik % 2])
                    km = (kv - qv) % 2
                    ikq = km[0] * 4 + km[1] * 2 + km[2]
                    def G(k, iw):
                        a = -inp.hk[:, :, k].copy()
                        for i in range(n):
                            a[i, i] = 1j * nu[iw + off] + inp.mu - inp.hk[i, i, k] - inp.dc[0, i] - inp.siw[iw + off, i]
                        return np.linalg.inv(a)
                    g1, g2 = G(ik, iwf), G(ikq, iwf - iwb)
                    for i in range(n):
                        for j in range(n):
                            for l in range(n):
                                for k in range(n):
                                    acc[i, j, l, k] += -inp.beta * g1[i, l] * g2[k, j]
                ref = (acc / 8).reshape(n * n, n * n)
                assert rel_err(cq[:, :, d], ref) < 1e-13


def _flat_band(eps, mu, beta, iwmax):
    """One orbital with H(k) = eps on a 2x2x2 grid, Sigma = 0, dc = 0."""
    wl = synth.workload("sc1_k4", nk=(2, 2, 2), iwmax=iwmax)
    inp = synth.build_inputs(wl)
    nkp = 8
    return dataclasses.replace(inp, hk=np.full((1, 1, nkp), eps, dtype=np.complex128),
                               siw=np.zeros((2 * iwmax, 1), dtype=np.complex128),
                               dc=np.zeros((2, 1), dtype=np.complex128), mu=mu, beta=beta)


def test_one_particle_flat_band_closed_forms():
    """A flat band without self-energy: G(nu) = 1/(i nu + mu - eps) at every k, and the
    occupation 2/beta sum_{n<M} Re G + 1/2 tends to the Fermi function 1/(exp(beta(eps-mu)) + 1);
    the neglected tail is beta (mu - eps) / (2 pi^2 M) to first order."""
    eps, mu, beta, M = 0.2, 0.5, 10.0, 20000
    inp = _flat_band(eps, mu, beta, M)
    giw = adga_oracle.get_giw(inp)
    iw = np.arange(-M, M)
    nu = np.pi * (2 * iw + 1) / beta
    assert np.abs(giw[:, 0] - 1.0 / (1j * nu + mu - eps)).max() < 1e-15
    nf = adga_oracle.get_nfock(inp)
    fermi = 1.0 / (np.exp(beta * (eps - mu)) + 1.0)
    tail = beta * (mu - eps) / (2 * np.pi ** 2 * M)
    assert abs(nf[0, 0, 0].real + tail - fermi) < 1e-8
    assert np.abs(nf.imag).max() == 0.0
    # the same with the self-energy passed as sigma_sum = 0 inside the box
    ss = np.zeros((1, 1, inp.wl.nf, 8), dtype=np.complex128)
    ndk, nd = adga_oracle.get_ndga(inp, ss)
    assert rel_err(ndk, nf) < 1e-14 and abs(nd[0] - nf[0, 0, 0]) < 1e-14
    sl, gl = adga_oracle.get_sigma_g_loc(inp, ss)
    assert np.abs(sl).max() == 0.0 and rel_err(gl[:, 0, 0], giw[:, 0]) < 1e-14


def test_one_particle_constant_self_energy_shifts_the_band():
    """A k- and nu-independent sigma_sum (inside the box) and siw (outside) equal to s act like
    eps -> eps + s: get_ndga and get_sigma_g_loc equal the self-energy-free routines at the
    shifted band, and a diagonal dc does the same with the opposite sign of mu."""
    eps, mu, beta, M, s = -0.3, 0.1, 5.0, 400, 0.25
    base = _flat_band(eps, mu, beta, M)
    shifted = _flat_band(eps + s, mu, beta, M)
    inp = dataclasses.replace(base, siw=np.full((2 * M, 1), s, dtype=np.complex128))
    ss = np.full((1, 1, inp.wl.nf, 8), s, dtype=np.complex128)
    ndk, _ = adga_oracle.get_ndga(inp, ss)
    assert rel_err(ndk, adga_oracle.get_nfock(shifted)) < 1e-13
    _, gl = adga_oracle.get_sigma_g_loc(inp, ss)
    assert rel_err(gl[:, 0, 0], adga_oracle.get_giw(shifted)[:, 0]) < 1e-13
    dcin = dataclasses.replace(base, dc=np.array([[s], [0.0]], dtype=np.complex128))
    assert rel_err(adga_oracle.get_nfock(dcin), adga_oracle.get_nfock(shifted)) < 1e-13


def _local_limit_inputs():
    wl = synth.workload("two_k4", nk=(1, 1, 1))
    inp = synth.build_inputs(wl)
    inp.hk[0, 1, 0] = inp.hk[1, 0, 0] = 0.0
    hp = adga_oracle.HotPath(inp)
    for iw in range(-inp.iwmax, inp.iwmax):
        inp.giw[iw + inp.iwmax] = np.diag(hp.g_matrix(np.array([0]), iw)[0])
    return wl, inp


def test_finalize_local_limit_is_the_double_frequency_sum():
    """One k point, orbital-diagonal H: chi_nl = 0 and bubble_nl = 0, so the assembled chi_r(q, omega)
    (main.F90:935-936, 1050-1063, 1127-1140) is the local susceptibility, which is the double frequency
    sum beta^-2 sum_{nu nu'} chi_loc,r({i nu}, {l nu'}) of the (transpose-symmetrised) generalized
    susceptibility; Sigma_nl = 0, so Sigma_DGA is the DMFT self-energy on the diagonal.  A dropped
    chi_loc or bubble_loc term, or chi_loc taken from the wrong omega, fails it."""
    wl, inp = _local_limit_inputs()
    hp = adga_oracle.HotPath(inp)
    pts = [(w, 0) for w in range(-wl.iwbmax_small, wl.iwbmax_small + 1)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    out = hp.run(pts, cl)
    fin = hp.finalize(out, pts)
    n2, nf = wl.n ** 2, wl.nf
    for p, (w, _) in enumerate(pts):
        for r, ch in (("d", "dens"), ("m", "magn")):
            x = cl(w, ch)
            x = 0.5 * (x + x.T)
            ref = np.asarray(x).reshape(nf, n2, nf, n2).sum(axis=(0, 2)) / inp.beta ** 2
            assert rel_err(fin["chi_" + r][:, :, p], ref) < 1e-12, (w, r)
    siw = inp.siw[inp.iwmax - wl.iwfmax_small:inp.iwmax + wl.iwfmax_small]
    for a in range(wl.n):
        assert rel_err(fin["sigma_dga"][a, a, :, 0], siw[:, a]) < 1e-14
    assert abs(fin["sigma_dga"][0, 1]).max() < 1e-14


def test_finalize_F_zero_is_the_full_bubble():
    """chi_loc = chi0^w: chi_nl = 0 and chi_loc(omega) = 0, so the assembled chi_r(q, omega) is the
    full bubble bubble_nl + bubble_loc = beta^-2 sum_nu chi0^q(nu) in both channels, and Sigma_DGA =
    siw on the diagonal."""
    inp = _inp("t2g_k4", noise=0.0)
    wl = inp.wl
    hp = adga_oracle.HotPath(inp)
    pts = [(w, q) for w in (-1, 0, 2) for q in range(0, 64, 7)]
    out = hp.run(pts, lambda w, c: synth.chi_loc(inp, w, c))
    fin = hp.finalize(out, pts)
    for p, (w, q) in enumerate(pts):
        ref = hp.chi0q(q, w).sum(axis=2) / inp.beta ** 2
        assert rel_err(fin["chi_d"][:, :, p], ref) < 1e-12
        assert rel_err(fin["chi_m"][:, :, p], ref) < 1e-12
    siw = inp.siw[inp.iwmax - wl.iwfmax_small:inp.iwmax + wl.iwfmax_small]
    for a in range(wl.n):
        assert np.abs(fin["sigma_dga"][a, a] - siw[:, a][:, None]).max() < 1e-14


def test_vq_zero_equals_no_vq():
    """V(q) = 0 given explicitly must reproduce the run without V(q) (every V term is linear in V),
    and a q-independent V gives a k-independent Fock term."""
    inp = _inp("t2g_k4_vq")
    pts = [(0, 0), (0, 3), (1, 5)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    zero = dataclasses.replace(inp, v_q=np.zeros_like(inp.v_q))
    none = dataclasses.replace(inp, v_q=None)
    a = adga_oracle.HotPath(zero).run(pts, cl)
    b = adga_oracle.HotPath(none).run(pts, cl)
    for k in ("chi_d", "chi_m", "sigma_nl", "sigma_dmft"):
        assert rel_err(a[k], b[k]) < 1e-14, k
    assert np.abs(a["sigma_hf"]).max() == 0.0
