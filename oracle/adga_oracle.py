"""CPU oracle (numpy, complex128) of ADGA's per-(q,omega) hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP path: only
`tests/`, `__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py`
may import it, and never as the thing that is measured or shipped.

It restates the reference's algorithm in numpy, keeping the reference's array
shapes and index meaning (first index fastest, 0-based here):

* G(k,nu)                   one_particle_quant_module.f90:81-104, 240-271; aux.f90:288-325
* k - q                     kq_tools.f90:96-108
* chi0^q bubble             one_particle_quant_module.f90:228-290; main.F90:644-657
* chi0^w and its inverse    one_particle_quant_module.f90:192-225; main.F90:509-522
* local vertex preparation  main.F90:525-625 (chi_loc from read_vertex, hdf5_module.f90:708-770)
* BSE build                 main.F90:673-720
* BSE inversion             main.F90:726-743 -> lapack_module.f90:40-70 (zgetrf + zgetri)
* eta                       main.F90:749-757
* chi contraction           susc_module.f90:34-56; main.F90:573-577, 661-668, 766-773
* EOM (dynamic)             eom_module.f90:165-233
* EOM (local, static)       eom_module.f90:39-161; main.F90:635-639
* V(q) branch               main.F90:497-503 (read_vq per point), 710-719 (density BSE term);
                            eom_module.f90:72-108 (Hartree/Fock), 183-207 (V terms of m_tot)
* Sigma assembly            eom_module.f90:237-251; main.F90:840-855
* chi assembly              main.F90:930-936, 1047-1063, 1124-1140
* one-particle routines     one_particle_quant_module.f90:40-79 (get_giw), 124-180
                            (get_sigma_g_loc), 337-361 (get_nfock), 363-393 (get_ndga)
* large-range bubble        config_module.f90:467-477 (debug option "Bubble": iwstart/iwstop),
                            main.F90:513-520, 644-663
* Matsubara tail            no reference counterpart: the local bubble summed beyond the large
                            range with the asymptotic expansion of giw (`tail_sums`,
                            `bubble_loc_tail`)
* lambda correction         no reference counterpart: the physical susceptibility of the
                            reduced output (hdf5_module.f90:1479-1495, chi_(ii),(jj) summed over
                            i, j) corrected as 1/(1/chi + lambda), lambda from the sum rule
                            (`chi_phys`, `lambda_sums`, `lambda_solve`)

Parity pins (tests/test_oracle_*.py): golden vectors written by the reference's
own Fortran modules (oracle/_ref, built from the reference sources with
amdflang + MKL) on small workloads (also with a reordered EOM k list, a q path
and V(q)), and closed forms: the F = 0 limit, the local limit, a brute-force
loop restatement on tiny grids, numpy inverse for the adjugate, the Matsubara
symmetries; `finalize` (the chi and Sigma assembly after the loop, which the
reference does inline in main.F90 and no golden vector covers) by the local
limit (chi = beta^-2 sum_{nu nu'} chi_loc) and the F = 0 limit (chi = the full
bubble, Sigma_DGA = siw); the Matsubara-tail sums by high-precision summation
(mpmath) and the tail-corrected local bubble of a flat band by its closed form
(the Lindhard function [f(a) - f(b)] / (a - b - i omega)).
"""
from __future__ import annotations

import numpy as np

__all__ = ["HotPath", "inverse_small", "kmq_table", "get_giw", "get_nfock", "get_sigma_g_loc", "get_ndga",
           "tail_sums", "giw_moments", "bubble_loc_tail", "chi_phys", "lambda_sums", "lambda_solve"]


def inverse_small(a):
    """Batched inverse of (..., n, n) matrices: closed-form adjugate for n <= 3
    (aux.f90:288-325, main routine's choice at one_particle_quant_module.f90:245-256),
    LAPACK otherwise (lapack_module.f90:40-70)."""
    n = a.shape[-1]
    if n == 1:
        return 1.0 / a
    if n == 2:
        d = 1.0 / (a[..., 0, 0] * a[..., 1, 1] - a[..., 0, 1] * a[..., 1, 0])
        b = np.empty_like(a)
        b[..., 0, 0] = d * a[..., 1, 1]
        b[..., 1, 0] = -d * a[..., 1, 0]
        b[..., 0, 1] = -d * a[..., 0, 1]
        b[..., 1, 1] = d * a[..., 0, 0]
        return b
    if n == 3:
        A = lambda i, j: a[..., i - 1, j - 1]  # noqa: E731  (1-based like the formula)
        d = 1.0 / (A(1, 1) * A(2, 2) * A(3, 3) - A(1, 1) * A(2, 3) * A(3, 2)
                   - A(1, 2) * A(2, 1) * A(3, 3) + A(1, 2) * A(2, 3) * A(3, 1)
                   + A(1, 3) * A(2, 1) * A(3, 2) - A(1, 3) * A(2, 2) * A(3, 1))
        b = np.empty_like(a)
        b[..., 0, 0] = d * (A(2, 2) * A(3, 3) - A(2, 3) * A(3, 2))
        b[..., 1, 0] = -d * (A(2, 1) * A(3, 3) - A(2, 3) * A(3, 1))
        b[..., 2, 0] = d * (A(2, 1) * A(3, 2) - A(2, 2) * A(3, 1))
        b[..., 0, 1] = -d * (A(1, 2) * A(3, 3) - A(1, 3) * A(3, 2))
        b[..., 1, 1] = d * (A(1, 1) * A(3, 3) - A(1, 3) * A(3, 1))
        b[..., 2, 1] = -d * (A(1, 1) * A(3, 2) - A(1, 2) * A(3, 1))
        b[..., 0, 2] = d * (A(1, 2) * A(2, 3) - A(1, 3) * A(2, 2))
        b[..., 1, 2] = -d * (A(1, 1) * A(2, 3) - A(1, 3) * A(2, 1))
        b[..., 2, 2] = d * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1))
        return b
    return np.linalg.inv(a)


def kmq_table(nk, q_data):
    """kq_ind(nkp, nqp): 0-based index of k - q for every k and every q of the list
    (index_kq, kq_tools.f90:59-71 with k_minus_q 96-108)."""
    nkx, nky, nkz = nk
    ik = np.arange(nkx * nky * nkz)
    iz, iy, ix = ik % nkz, (ik // nkz) % nky, ik // (nky * nkz)
    q = np.asarray(q_data) - 1
    lz, ly, lx = q % nkz, (q // nkz) % nky, q // (nky * nkz)
    return (((iz[:, None] - lz[None, :]) % nkz) + ((iy[:, None] - ly[None, :]) % nky) * nkz
            + ((ix[:, None] - lx[None, :]) % nkx) * nky * nkz)


# ---------------------------------------------------------------- one-particle routines
def _ginv_1p(inp, ks, iw, sigma=None):
    """G^-1(k, nu_iw) for 0-based k indices `ks`, shape (len, n, n) with [.., a, b]: -H(k) - Sigma
    off the diagonal, i nu + mu - H_ii - dc(1, i) - Sigma_ii on it; Sigma = siw(iw, :) (diagonal)
    when `sigma` is None, else the (len, n, n) matrices given (get_gkiw 81-104, get_gkiw_dga
    107-121, get_sigma_g_loc 153-160)."""
    n = inp.wl.n
    h = np.asarray(inp.hk)[:, :, ks].transpose(2, 0, 1).astype(np.complex128)
    g = -h if sigma is None else -h - sigma
    nu = np.pi * (2 * iw + 1) / inp.beta
    d = np.arange(n)
    sd = inp.siw[iw + inp.iwmax, :][None, :] if sigma is None else sigma[:, d, d]
    g[:, d, d] = 1j * nu + inp.mu - h[:, d, d] - inp.dc[0, :][None, :] - sd
    return g


def get_giw(inp):
    """giw(-iwmax:iwmax-1, n) as a (2*iwmax, n) array: (1/nkp) sum_k diag G(k, nu) for nu >= 0 and
    giw(-nu-1) = giw(nu)^* (get_giw, one_particle_quant_module.f90:40-79; the reference inverts
    with inverse_matrix, i.e. LAPACK)."""
    nkp, off = inp.hk.shape[2], inp.iwmax
    ks = np.arange(nkp)
    out = np.zeros((2 * off, inp.wl.n), dtype=np.complex128)
    for iw in range(off):
        g = np.linalg.inv(_ginv_1p(inp, ks, iw))
        out[iw + off] = np.diagonal(g, axis1=1, axis2=2).sum(axis=0)
        out[-iw - 1 + off] = np.conj(out[iw + off])
    return out / nkp


def get_nfock(inp):
    """n_fock(nkp, n, n) = 2/beta sum_{iw<iwmax} Re G(k, nu) + 1/2 on the diagonal (get_nfock,
    one_particle_quant_module.f90:337-361; complex array with zero imaginary part)."""
    nkp, n = inp.hk.shape[2], inp.wl.n
    ks = np.arange(nkp)
    acc = np.zeros((nkp, n, n))
    for iw in range(inp.iwmax):
        acc += np.linalg.inv(_ginv_1p(inp, ks, iw)).real
    out = (2.0 * acc / inp.beta).astype(np.complex128)
    out[:, np.arange(n), np.arange(n)] += 0.5
    return out


def get_sigma_g_loc(inp, sigma_sum):
    """(sigma_loc, gloc) of get_sigma_g_loc (one_particle_quant_module.f90:124-180): sigma_sum
    (n, n, 2*iwfmax_small, nkp) (first index fastest, after add_siw_dmft); sigma_loc its k mean;
    gloc (2*iwmax, n, n) = (1/nkp) sum_k G with sigma_sum inside the box and siw outside, and
    gloc(-nu-1) = gloc(nu)^H."""
    nkp, off, iwf = inp.hk.shape[2], inp.iwmax, inp.wl.iwfmax_small
    ks = np.arange(nkp)
    sloc = sigma_sum.sum(axis=3) / nkp
    gloc = np.zeros((2 * off, inp.wl.n, inp.wl.n), dtype=np.complex128)
    for iw in range(off):
        sig = sigma_sum[:, :, iw + iwf, :].transpose(2, 0, 1) if iw < iwf else None
        gloc[iw + off] = np.linalg.inv(_ginv_1p(inp, ks, iw, sig)).sum(axis=0)
        gloc[-iw - 1 + off] = np.conj(gloc[iw + off]).T
    return sloc, gloc / nkp


def get_ndga(inp, sigma_sum):
    """(n_dga_k, n_dga) of get_ndga (one_particle_quant_module.f90:363-393): sigma_sum
    (n, n, 2*iwfmax_small, nkp_eom) on the EOM k list; n_dga_k(nkp_eom, n, n) = 2/beta sum_iw
    Re G(k_data_eom(ik), nu) + 1/2 on the diagonal with sigma_sum inside the box and the diagonal
    siw outside; n_dga the mean of its diagonal."""
    n, iwf = inp.wl.n, inp.wl.iwfmax_small
    ks = np.asarray(inp.k_data_eom) - 1
    acc = np.zeros((len(ks), n, n))
    for iw in range(inp.iwmax):
        sig = sigma_sum[:, :, iw + iwf, :].transpose(2, 0, 1) if iw < iwf else None
        acc += np.linalg.inv(_ginv_1p(inp, ks, iw, sig)).real
    ndk = (2.0 * acc / inp.beta).astype(np.complex128)
    ndk[:, np.arange(n), np.arange(n)] += 0.5
    nd = np.diagonal(ndk, axis1=1, axis2=2).sum(axis=0) / len(ks)
    return ndk, nd


def tail_sums(beta, K, m):
    """S[(p, r)] = sum over n outside [-K, K-1] of x^-p y^-r with x = i nu_n, y = i nu_{n-m} = x - i omega_m,
    for (p, r) in (1,1), (2,1), (1,2), (3,1), (2,2), (1,3): by partial fractions in w = x - y into sums of
    single powers, X_s = sum x^-s and Y_s = sum y^-s, whose even powers are Hurwitz zeta tails and whose
    odd powers cancel pairwise (nu_{-n-1} = -nu_n) except for 2|m| terms of the shifted set (the first
    power as the limit of symmetric truncations, the sums with p + r >= 2 converging absolutely)."""
    from scipy.special import zeta
    nu = lambda j: np.pi * (2 * np.asarray(j, dtype=float) + 1) / beta  # noqa: E731
    c = beta / (2 * np.pi)
    B = {2: 2 * c ** 2 * zeta(2, K + 0.5), 4: 2 * c ** 4 * zeta(4, K + 0.5)}   # sum nu^-s over the outside
    X = {1: 0.0, 2: (1j) ** -2 * B[2], 3: 0.0}
    if m == 0:
        return {(1, 1): X[2], (2, 1): 0.0, (1, 2): 0.0, (3, 1): B[4], (2, 2): B[4], (1, 3): B[4]}
    M = abs(m)
    inner = np.arange(K - M, K + M)                    # the 2|m| unpaired terms of an odd power
    lower, upper = np.arange(K - M, K), np.arange(K, K + M)
    A = {1: np.sign(m) * np.sum(1 / nu(inner)), 3: np.sign(m) * np.sum(1 / nu(inner) ** 3),
         2: B[2] + np.sum(1 / nu(lower) ** 2) - np.sum(1 / nu(upper) ** 2)}
    Y = {s: (1j) ** -s * A[s] for s in (1, 2, 3)}
    w = 2j * np.pi * m / beta
    D = Y[1] - X[1]
    return {(1, 1): D / w,
            (2, 1): D / w ** 2 - X[2] / w,
            (1, 2): Y[2] / w - D / w ** 2,
            (3, 1): D / w ** 3 - X[2] / w ** 2 - X[3] / w,
            (2, 2): (Y[2] + X[2]) / w ** 2 - 2 * D / w ** 3,
            (1, 3): Y[3] / w - Y[2] / w ** 2 + D / w ** 3}


def giw_moments(giw, beta, iwmax):
    """(c1, c2) of giw = 1/(i nu) + c1/(i nu)^2 + c2/(i nu)^3 + c3/(i nu)^4 + ..., per orbital, from the two
    highest positive frequencies: u = nu^2 Re g = -c1 + c3/nu^2 and v = nu^3 (Im g + 1/nu) = c2 - c4/nu^2
    are linear in h = 1/nu^2 up to O(h^2), so their values at the two frequencies give c1 and c2 with an
    error O(nu^-4)."""
    nu = np.pi * (2 * np.array([iwmax - 1, iwmax - 2]) + 1) / beta
    g = giw[[2 * iwmax - 1, 2 * iwmax - 2]]                   # (2, n)
    h = 1 / nu ** 2
    u = nu[:, None] ** 2 * g.real
    v = nu[:, None] ** 3 * (g.imag + 1 / nu[:, None])
    c1 = -(u[0] * h[1] - u[1] * h[0]) / (h[1] - h[0])
    c2 = (v[0] * h[1] - v[1] * h[0]) / (h[1] - h[0])
    return c1, c2


def bubble_loc_tail(giw, beta, iwmax, K, iwb):
    """The local bubble's frequency sum beyond the large range, nu outside [-K, K-1]:
    -1/beta sum g_i(nu) g_j(nu - omega) on the diagonal i1 = i*n + j, with g expanded to third order in
    1/(i nu) (the dropped terms are O(nu^-5), whose symmetric sum is O(K^-4) in units of beta/2pi)."""
    S = tail_sums(beta, K, iwb)
    c1, c2 = giw_moments(giw, beta, iwmax)
    n = len(c1)
    t = np.zeros(n * n, dtype=np.complex128)
    for i in range(n):
        for j in range(n):
            t[i * n + j] = -(S[1, 1] + c1[i] * S[2, 1] + c1[j] * S[1, 2] + c2[i] * S[3, 1]
                             + c1[i] * c1[j] * S[2, 2] + c2[j] * S[1, 3]) / beta
    return np.diag(t)


def chi_phys(chi):
    """sum_{i,j} chi[(i,i), (j,j)] of chi (n2, n2, ...): the reduced susceptibility's entries
    (hdf5_module.f90:1479-1495: i2 = i1, i4 = i3) summed over both orbitals."""
    n = int(round(np.sqrt(chi.shape[0])))
    d = np.arange(n) * n + np.arange(n)
    return chi[np.ix_(d, d)].sum(axis=(0, 1))


def lambda_sums(chi, beta, nqp, lam):
    """S(lambda) = 1/(beta nqp) sum over every point of chi^lambda, chi^lambda = 1/(1/chi + lambda), and
    dS/dlambda = -1/(beta nqp) sum (chi^lambda)^2, for the physical susceptibilities chi of the points."""
    cl = chi / (1 + lam * chi)
    return cl.sum() / (beta * nqp), -(cl * cl).sum() / (beta * nqp)


def lambda_solve(chi, beta, nqp, target, chi0=None):
    """lambda with Re S(lambda) = target by Brent's method (scipy.optimize.brentq) on (lambda_min, lambda_hi],
    lambda_min = -1/max Re chi(omega = 0) the pole of the largest omega = 0 susceptibility (`chi0`; all
    points if not given): Re S decreases monotonically from +inf there (for positive Re chi at omega = 0)."""
    from scipy.optimize import brentq
    c0 = chi if chi0 is None else chi0
    lmin = -1.0 / np.max(c0.real)
    f = lambda lam: lambda_sums(chi, beta, nqp, lam)[0].real - target  # noqa: E731
    hi = 1.0
    while f(hi) > 0:
        hi *= 2.0
    lo = lmin + 1e-12 * max(1.0, abs(lmin))
    return brentq(f, lo, hi, xtol=1e-15, rtol=1e-15, maxiter=500)


class HotPath:
    """The per-(q,omega) loop body of main.F90:486-799 over given inputs."""

    def __init__(self, inp, bse_inverse=None):
        self.inp = inp
        wl = inp.wl
        self.n = wl.n
        self.n2 = wl.n * wl.n
        self.iwf = wl.iwfmax_small
        self.iwb = wl.iwbmax_small
        self.nf = 2 * wl.iwfmax_small
        self.N = self.n2 * self.nf
        self.nkp = inp.hk.shape[2]
        self.nqp = len(inp.q_data)
        self.nkp_eom = len(inp.k_data_eom)
        self.beta = inp.beta
        self.off = inp.iwmax
        self.kq = kmq_table(wl.nk, inp.q_data)
        # bubble sums over nu in [lo, hi]: the BSE box, or with the debug option "Bubble" every frequency
        # for which G(nu - omega) exists (iwstart/iwstop, config_module.f90:467-477)
        self.bubble_range = getattr(wl, "bubble_range", 0)
        if self.bubble_range:
            self.lo, self.hi = -inp.iwmax + self.iwb, inp.iwmax - self.iwb - 1
        else:
            self.lo, self.hi = -self.iwf, self.iwf - 1
        # frequencies used by G(k, nu) and G(k-q, nu-omega)
        self.glo = self.lo - self.iwb
        self.ghi = self.hi + self.iwb
        self._gtab = None
        self.bse_inverse = bse_inverse or np.linalg.inv

    # ---------------------------------------------------------------- A1: G(k, nu)
    def g_matrix(self, iks, iw):
        """G(k, nu_iw) for 0-based k indices `iks`, shape (len, n, n) with [.., a, b]
        = G_ab (one_particle_quant_module.f90:240-271, get_gkiw 81-104)."""
        inp = self.inp
        n = self.n
        hk = inp.hk[:, :, iks]  # (n, n, K)
        a = -np.transpose(hk, (2, 0, 1)).copy()
        nu = inp.iw_data[iw + self.off]
        for i in range(n):
            a[:, i, i] = 1j * nu + inp.mu - hk[i, i, :] - inp.dc[0, i] - inp.siw[iw + self.off, i]
        return inverse_small(a)

    @property
    def gtab(self):
        """G table (n_w, nkp, n, n) over nu in [glo, ghi]."""
        if self._gtab is None:
            allk = np.arange(self.nkp)
            self._gtab = np.stack([self.g_matrix(allk, w) for w in range(self.glo, self.ghi + 1)])
        return self._gtab

    def G(self, iw):
        return self.gtab[iw - self.glo]

    # ---------------------------------------------------------------- A4
    def chi0_loc(self, iwf, iwb):
        """Diagonal of chi0^w(nu) in compound index i1 = i*n + j (get_chi0_loc 192-207)."""
        g = self.inp.giw
        gi = g[iwf + self.off]
        gj = g[iwf - iwb + self.off]
        return (-self.beta * gi[:, None] * gj[None, :]).reshape(-1)

    def chi0_loc_inv(self, iwf, iwb):
        """Diagonal of [chi0^w(nu)]^-1 (get_chi0_loc_inv 210-225)."""
        g = self.inp.giw
        gi = g[iwf + self.off]
        gj = g[iwf - iwb + self.off]
        return (-1.0 / (self.beta * gi[:, None] * gj[None, :])).reshape(-1)

    # ---------------------------------------------------------------- A3
    def chi0q(self, iq, iwb):
        """k-summed bubble chi0^q(nu) for every nu of the box, shape (n2, n2, Nf):
        chi0[(i,j),(l,k)] = -beta/nkp sum_k G(k,nu)_il G(k-q,nu-omega)_kj
        (accumulate_chi0 one_particle_quant_module.f90:274-288, main.F90:647-657)."""
        n = self.n
        kq = self.kq[:, iq]
        out = np.empty((self.n2, self.n2, self.nf), dtype=np.complex128)
        for d in range(self.nf):
            iwf = d - self.iwf
            g1 = self.G(iwf)            # (K, n, n) [K, i, l]
            g2 = self.G(iwf - iwb)[kq]  # (K, n, n) [K, k, j]
            c = -self.beta * np.einsum("Kil,Kkj->ijlk", g1, g2, optimize=True)
            out[:, :, d] = c.reshape(self.n2, self.n2) / self.nkp
        return out

    def chi0q_outer(self, iq, iwb):
        """sum of chi0^q(nu) over the large range outside the box (main.F90:644-655 with the debug option
        "Bubble": those frequencies enter bubble_nl only)."""
        kq = self.kq[:, iq]
        acc = np.zeros((self.n2, self.n2), dtype=np.complex128)
        for iwf in range(self.lo, self.hi + 1):
            if -self.iwf <= iwf < self.iwf:
                continue
            c = -self.beta * np.einsum("Kil,Kkj->ijlk", self.G(iwf), self.G(iwf - iwb)[kq], optimize=True)
            acc += c.reshape(self.n2, self.n2) / self.nkp
        return acc

    # ---------------------------------------------------------------- A9
    def calc_chi_qw(self, x, y):
        """beta^-2 sum_{i3,nu} X(:,{i3,nu}) Y(i3,:,nu) (susc_module.f90:34-56)."""
        yy = np.transpose(y, (2, 0, 1)).reshape(self.N, self.n2)
        return (x @ yy) / self.beta ** 2

    # ---------------------------------------------------------------- A4 + A5
    def local_omega(self, iwb, chi_loc_d, chi_loc_m):
        """Quantities updated when omega changes (main.F90:506-625)."""
        n2, nf, N = self.n2, self.nf, self.N
        winv = np.stack([self.chi0_loc_inv(d - self.iwf, iwb) for d in range(nf)], axis=1)  # (n2, nf)
        w = np.stack([self.chi0_loc(d - self.iwf, iwb) for d in range(nf)], axis=1)
        chi0w = np.zeros((n2, n2, nf), dtype=np.complex128)
        for d in range(nf):
            chi0w[np.arange(n2), np.arange(n2), d] = w[:, d]
        big = sum(self.chi0_loc(iwf, iwb) for iwf in range(self.lo, self.hi + 1))   # main.F90:513-520
        bubble_loc_tmp = np.diag(big) / self.beta ** 2
        out = {"winv": winv, "chi0w": chi0w, "bubble_loc_tmp": bubble_loc_tmp}
        for r, cl in (("d", chi_loc_d), ("m", chi_loc_m)):
            a = 0.5 * (cl + cl.T)                        # main.F90:529-537
            a = a * winv.T.reshape(-1)[None, :]          # column scaling, main.F90:542-553
            a[np.arange(N), np.arange(N)] -= 1.0
            gam = a.reshape(nf, n2, N).sum(axis=0)       # main.F90:557-571
            one = gam.copy()                             # main.F90:616-625
            for d in range(nf):
                one[np.arange(n2), d * n2 + np.arange(n2)] += 1.0
            out["chi0wF_" + r] = a
            out["gammaw_" + r] = gam
            out["oneplusgammaw_" + r] = one
            out["chi_loc_" + r] = self.calc_chi_qw(gam, chi0w)  # main.F90:573-577
        return out

    # ---------------------------------------------------------------- A6 - A9 for one point
    def point(self, iq, iwb, loc, chi0q=None):
        n2, nf, N = self.n2, self.nf, self.N
        if chi0q is None:
            chi0q = self.chi0q(iq, iwb)
        res = {"chi0q": chi0q}
        bsum = chi0q.sum(axis=2)
        if self.bubble_range:
            bsum = bsum + self.chi0q_outer(iq, iwb)
        bubble_nl = bsum / self.beta ** 2 - loc["bubble_loc_tmp"]   # main.F90:663
        chi0nl = chi0q - loc["chi0w"]
        res["bubble_nl"] = bubble_nl
        chi = {r: self.calc_chi_qw(loc["gammaw_" + r], chi0nl) for r in ("d", "m")}  # main.F90:666-667
        big = {r: np.zeros((N, N), dtype=np.complex128) for r in ("d", "m")}
        gammaqd = np.zeros((n2, N), dtype=np.complex128)
        v = self.v(iq)
        for d in range(nf):                                   # main.F90:679-720
            s = chi0q[:, :, d] * loc["winv"][:, d][None, :]
            s[np.arange(n2), np.arange(n2)] -= 1.0
            rows = slice(d * n2, (d + 1) * n2)
            for r in ("d", "m"):
                rect = s @ loc["chi0wF_" + r][rows, :]
                big[r][rows, :] = -rect
                if r == "d":
                    gammaqd += rect
            if v is not None:                                 # main.F90:710-719
                big["d"][rows, :] -= (2.0 / self.beta ** 2 * (chi0q[:, :, d] @ v)) @ loc["oneplusgammaw_d"]
        res["gammaqd"] = gammaqd
        for r in ("d", "m"):
            a = big[r] + np.eye(N)                            # main.F90:726-729
            res["A_" + r] = a
            ainv = self.bse_inverse(a) - np.eye(N)           # main.F90:734-743
            eta = loc["oneplusgammaw_" + r] @ ainv            # main.F90:749-757
            res["eta_" + r] = eta
            chi[r] = chi[r] + self.calc_chi_qw(eta, chi0q)    # main.F90:766-769
            res["chi_" + r] = chi[r]
        return res

    def v(self, iq):
        """V(q) of q list position iq (0-based) as read_vq returns it, or None without V(q)."""
        return None if self.inp.v_q is None else self.inp.v_q[:, :, iq]

    # ---------------------------------------------------------------- A10
    def mtot(self, etaqd, etaqm, gammaqd, v=None, gammawd=None):
        """m_tot of calc_eom_dynamic (eom_module.f90:183-207); with V(q) the density eta term takes
        U + V - U~/2 and -V gamma^w_d / (beta nqp) is added."""
        inp = self.inp
        f = 1.0 / (self.beta * self.nqp)
        uw = inp.u - 0.5 * inp.u_tilde + (0.0 if v is None else v)
        m = (-f) * (uw @ etaqd) + 1.5 * f * (inp.u_tilde @ etaqm) + f * (inp.u @ gammaqd)
        if v is not None:
            m = m - f * (v @ gammawd)
        return m

    def eom_dynamic(self, etaqd, etaqm, gammaqd, iq, iwb, sigma_nl, gammawd=None):
        """Sigma_nl(i,l,nu,k) += sum_{j,k'} m_tot({ji},{k'l,nu}) G(k-q,nu-omega)_{jk'}
        for nu >= 0, then Sigma(-nu-1) = Sigma(nu)^H (eom_module.f90:211-231)."""
        n, n2 = self.n, self.n2
        m = self.mtot(etaqd, etaqm, gammaqd, self.v(iq), gammawd)
        kidx = self.inp.k_data_eom - 1
        for d in range(self.iwf):                 # iwf = 0 .. iwfmax_small-1
            iwf = d
            col0 = n2 * (self.iwf + iwf)
            mm = m[:, col0:col0 + n2].reshape(n, n, n, n)  # [i, j, l, k']
            g = self.G(iwf - iwb)[self.kq[kidx, iq]]        # (K, j, k')
            add = np.einsum("ijlk,Kjk->ilK", mm, g, optimize=True)
            sigma_nl[:, :, self.iwf + iwf, :] += add
            sigma_nl[:, :, self.iwf - iwf - 1, :] = np.conj(np.transpose(sigma_nl[:, :, self.iwf + iwf, :], (1, 0, 2)))
        return m

    # ---------------------------------------------------------------- A11
    def eom_dmft(self, gammawd, iwb, sigma_dmft):
        """Local EOM with the DMFT G (eom_module.f90:115-161, default sig_ph branch)."""
        n, n2 = self.n, self.n2
        m = (-1.0 / self.beta) * (self.inp.u @ gammawd)
        for iwf in range(self.iwf):
            col0 = n2 * (self.iwf + iwf)
            # m_tot(i1 = i*n + k, i2 = col0 + l*n + k): j == k because g is diagonal
            mm = m[:, col0:col0 + n2].reshape(n, n, n, n)  # [i, k, l, k']
            gl = self.inp.giw[iwf - iwb + self.off]
            sigma_dmft[:, :, self.iwf + iwf] += np.einsum("iklk,k->il", mm, gl)
            sigma_dmft[:, :, self.iwf - iwf - 1] = np.conj(sigma_dmft[:, :, self.iwf + iwf].T)
        return m

    def eom_static_local(self, sigma_dmft):
        """Local Hartree term sum_j (2 u({ii},{jj}) - u({ji},{ji})) n_j (eom_module.f90:58-70)."""
        n = self.n
        u = self.inp.u
        for i in range(n):
            acc = 0.0
            for j in range(n):
                acc = acc + (2.0 * u[i * n + i, j * n + j] - u[i * n + j, i * n + j]) * self.inp.n_dmft[j]
            sigma_dmft[i, i, :] += acc

    @property
    def n_fock(self):
        """n_fock(nkp, n, n) of get_nfock (the Fock term's input, main.F90:271)."""
        if getattr(self, "_nfock", None) is None:
            self._nfock = get_nfock(self.inp)
        return self._nfock

    def eom_static_vq(self, iq, sigma_hf):
        """The V(q) part of calc_eom_static (eom_module.f90:72-108): at q = Gamma (iq == 0) the Hartree
        term 2 sum_j v({ii},{jj}) n_dmft(j) on the diagonal of every k; at every q the Fock term
        sigma_hf(i,l,k) -= sum_{j,k'} v({ik'},{lj}) / nqp * n_fock(k-q, j, k')."""
        n = self.n
        v = self.v(iq)
        if iq == 0:
            for i in range(n):
                for j in range(n):
                    sigma_hf[i, i, :] += 2.0 * v[i * n + i, j * n + j] * self.inp.n_dmft[j]
        kidx = self.inp.k_data_eom - 1
        nf = self.n_fock[self.kq[kidx, iq]]                   # (K, j, k')
        vv = v.reshape(n, n, n, n)                            # [i, k', l, j]
        sigma_hf -= np.einsum("iklj,Kjk->ilK", vv, nf) / self.nqp

    # ---------------------------------------------------------------- the loop
    def run(self, points, chi_loc_fn, keep=()):
        """Run the loop body for `points` = [(iwb, iq0), ...] in the given order
        (iq0 0-based into q_data).  `chi_loc_fn(iwb, channel)` supplies what
        read_vertex would return.  Returns per-point and accumulated outputs."""
        n, n2, nf = self.n, self.n2, self.nf
        npts = len(points)
        out = {
            "chi_d": np.zeros((n2, n2, npts), dtype=np.complex128),
            "chi_m": np.zeros((n2, n2, npts), dtype=np.complex128),
            "bubble_nl": np.zeros((n2, n2, npts), dtype=np.complex128),
            "sigma_nl": np.zeros((n, n, nf, self.nkp_eom), dtype=np.complex128),
            "sigma_dmft": np.zeros((n, n, nf), dtype=np.complex128),
            "sigma_hf": np.zeros((n, n, self.nkp_eom), dtype=np.complex128),
            "chi_loc_d": {}, "chi_loc_m": {}, "bubble_loc": {},
            "points": {},
        }
        do_eom = self.inp.wl.do_eom
        cur = None
        loc = None
        for p, (iwb, iq) in enumerate(points):
            if iwb != cur:
                cur = iwb
                loc = self.local_omega(iwb, chi_loc_fn(iwb, "dens"), chi_loc_fn(iwb, "magn"))
                if iq == 0:
                    out["bubble_loc"][iwb] = loc["bubble_loc_tmp"]
                    if self.bubble_range == 2:
                        out["bubble_loc"][iwb] = out["bubble_loc"][iwb] + bubble_loc_tail(
                            self.inp.giw, self.beta, self.off, self.hi + 1, iwb)
                    out["chi_loc_d"][iwb] = loc["chi_loc_d"]
                    out["chi_loc_m"][iwb] = loc["chi_loc_m"]
            if do_eom and iwb == 0 and iq == 0:                    # main.F90:635-637
                self.eom_static_local(out["sigma_dmft"])
            if do_eom and iwb == 0 and self.inp.v_q is not None:
                self.eom_static_vq(iq, out["sigma_hf"])
            if do_eom and iq == 0:                                 # main.F90:639
                self.eom_dmft(loc["gammaw_d"], iwb, out["sigma_dmft"])
            res = self.point(iq, iwb, loc)
            out["chi_d"][:, :, p] = res["chi_d"]
            out["chi_m"][:, :, p] = res["chi_m"]
            out["bubble_nl"][:, :, p] = res["bubble_nl"]
            if do_eom:
                self.eom_dynamic(res["eta_d"], res["eta_m"], res["gammaqd"], iq, iwb, out["sigma_nl"],
                                 loc["gammaw_d"])
            if (iwb, iq) in keep:
                out["points"][(iwb, iq)] = res
        return out

    def finalize(self, out, points):
        """chi_r(q,omega) = chi_nl + bubble_nl + chi_loc(omega) with chi_loc = calc_chi_qw(gamma, chi0w)
        + bubble_loc (main.F90:935-936, 1047-1063, 1124-1140); Sigma_DGA = Sigma_nl + Sigma_DMFT
        (add_siw_dmft, eom_module.f90:237-251)."""
        fin = {}
        for r in ("d", "m"):
            tot = out["chi_" + r] + out["bubble_nl"]
            for p, (iwb, iq) in enumerate(points):
                if iwb in out["chi_loc_" + r]:
                    tot[:, :, p] += out["chi_loc_" + r][iwb] + out["bubble_loc"][iwb]
            fin["chi_" + r] = tot
        s = out["sigma_nl"].copy()
        for a in range(self.n):
            s[a, a, :, :] += self.inp.siw[self.off - self.iwf:self.off + self.iwf, a][:, None]
        fin["sigma_dga"] = s
        return fin
