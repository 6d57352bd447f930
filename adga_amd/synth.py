"""Seeded synthetic inputs for the per-(q,omega) hot path.

The reference reads its inputs from HDF5 files (H(k), Sigma_DMFT, G_DMFT, the
local two-particle susceptibility) and a config file.  None of the SrVO3 input
files ship with the reference, so every workload here is synthetic and fully
determined by a `Workload` description and its seed:

* H(k): an analytic tight-binding model on the k-grid ordered like
  `k_index_3` (kq_tools.f90:119-125: ik = 1 + kz + ky*nkz + kx*nky*nkz).
* Sigma_DMFT(nu): an orbital-diagonal Hubbard-I-like form
  h_a + s_a^2/(i nu - e_a), which satisfies Sigma(-nu-1) = Sigma(nu)^*.
* g(nu): the k-summed diagonal of G, computed like `get_giw`
  (one_particle_quant_module.f90:40-78).
* n_dmft: like `get_ndmft` (one_particle_quant_module.f90:295-313).
* U, U~: Kanamori matrices in compound index, like `create_u`
  (interaction_module.f90:182-291) for one correlated atom.
* chi_loc_r(omega): what `read_vertex` (hdf5_module.f90:629-773) returns, i.e.
  the local bubble -beta g_i(nu) g_j(nu-omega) on the diagonal plus a seeded
  Hermitian complex noise 1e-3 (X + X^H)/2.

All arrays are complex128 and carry the reference's Fortran shapes and index
meaning (first index fastest): hk(n,n,nkp), siw/giw(2*iwmax, n) with frequency
offset iwmax, u(n^2,n^2), chi_loc(N,N).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

__all__ = ["Workload", "Inputs", "WORKLOADS", "workload", "build_inputs", "chi_loc",
           "kanamori_u", "generate_q_vol", "k_minus_q", "PARAMS", "Tangents", "tangents", "dchi_loc"]


@dataclasses.dataclass(frozen=True)
class Workload:
    name: str
    n: int                      # orbitals (ndim)
    nk: tuple                   # (nkx, nky, nkz)
    iwfmax_small: int           # Nf/2
    iwbmax_small: int           # N_omega = 2*iwbmax_small + 1
    beta: float
    U: float
    J: float
    model: str                  # "sc1", "t2g", "rand5"
    t: float = 0.0
    tp: float = 0.0
    seed: int = 20261202
    nq: tuple | None = None     # q grid (defaults to the k grid)
    iwmax: int | None = None    # 1P frequencies (defaults to 4*(Nf/2+Nw/2))
    noise: float = 1e-3
    k_eom: str = "all"          # EOM k list: "all" (the grid in order) or "sub" (every third k, reversed)
    q_list: str = "vol"         # q list: "vol" (generate_q_vol on qgrid) or "path" (a path through the grid)
    vq: float = 0.0             # scale of the synthetic non-local interaction V(q) (0: no V(q))
    bubble_range: int = 0       # bubble sums over the BSE box (0), over every frequency the 1P data
                                # allow (1, the reference's debug option "Bubble"), or 1 + the
                                # asymptotic Matsubara tail beyond them (2)

    @property
    def nkp(self):
        return self.nk[0] * self.nk[1] * self.nk[2]

    @property
    def qgrid(self):
        return self.nq if self.nq is not None else self.nk

    @property
    def nqp(self):
        if self.q_list == "path":
            return len(q_path(self.nk))
        q = self.qgrid
        return q[0] * q[1] * q[2]

    @property
    def do_eom(self):
        # the reference stops when a q path is combined with the EOM (main.F90:341-347)
        return self.q_list != "path"

    @property
    def nf(self):
        return 2 * self.iwfmax_small

    @property
    def maxdim(self):
        return self.n * self.n * self.nf

    @property
    def n_iwmax(self):
        if self.iwmax is not None:
            return self.iwmax
        return 4 * (self.iwfmax_small + self.iwbmax_small)


WORKLOADS = {
    # 1-orbital Hubbard model, simple cubic, NN t=0.25, U=2, beta=10, half filling.
    "sc1_k8": Workload("sc1_k8", 1, (8, 8, 8), 20, 5, 10.0, 2.0, 0.0, "sc1", t=0.25, seed=20261202),
    # 3-orbital t2g cubic model, NN t=-0.26, NNN t'=-0.08, Kanamori U=5, J=0.75.
    "t2g_k20": Workload("t2g_k20", 3, (20, 20, 20), 60, 15, 10.0, 5.0, 0.75, "t2g",
                        t=-0.26, tp=-0.08, seed=20261203),
    # 5-orbital d-shell, random symmetric hopping to 3rd neighbours (scale 0.3).
    "d5_k16": Workload("d5_k16", 5, (16, 16, 16), 80, 20, 10.0, 4.0, 0.7, "rand5",
                       t=0.3, seed=20261205),
    # 3-orbital t2g model on a 12^3 grid with 2x40 x 21 frequencies.
    "t2g_k12": Workload("t2g_k12", 3, (12, 12, 12), 40, 10, 10.0, 5.0, 0.75, "t2g",
                        t=-0.26, tp=-0.08, seed=20261206),
    # Small cases used by the parity tests (each finishes in seconds on the CPU).
    # EOM on a reordered third of the k grid with q on a 2x2x2 sub-grid (k_path_eom + q volume)
    "t2g_k4_keom": Workload("t2g_k4_keom", 3, (4, 4, 4), 6, 2, 10.0, 5.0, 0.75, "t2g",
                            t=-0.26, tp=-0.08, seed=16, nq=(2, 2, 2), k_eom="sub"),
    # susceptibility along a q path (no EOM, as the reference requires)
    "t2g_k6_qpath": Workload("t2g_k6_qpath", 3, (6, 6, 6), 6, 2, 10.0, 5.0, 0.75, "t2g",
                             t=-0.26, tp=-0.08, seed=17, q_list="path"),
    # non-local interaction V(q): density-channel BSE term, Hartree/Fock, V terms of the EOM
    "two_k4_vq": Workload("two_k4_vq", 2, (4, 4, 2), 4, 2, 8.0, 3.0, 0.5, "rand5",
                          t=0.3, seed=18, vq=0.4),
    "t2g_k4_vq": Workload("t2g_k4_vq", 3, (4, 4, 4), 6, 2, 10.0, 5.0, 0.75, "t2g",
                          t=-0.26, tp=-0.08, seed=19, vq=0.5, k_eom="sub"),
    # bubble_loc and bubble_nl over the large frequency range (debug option "Bubble")
    "t2g_k4_bub": Workload("t2g_k4_bub", 3, (4, 4, 4), 6, 2, 10.0, 5.0, 0.75, "t2g",
                           t=-0.26, tp=-0.08, seed=20, iwmax=20, bubble_range=1),
    "two_k4_bub": Workload("two_k4_bub", 2, (4, 4, 2), 4, 2, 8.0, 3.0, 0.5, "rand5",
                           t=0.3, seed=21, iwmax=15, bubble_range=1, k_eom="sub"),
    "t2g_k4": Workload("t2g_k4", 3, (4, 4, 4), 6, 2, 10.0, 5.0, 0.75, "t2g",
                       t=-0.26, tp=-0.08, seed=11),
    "t2g_k6": Workload("t2g_k6", 3, (6, 6, 6), 12, 3, 10.0, 5.0, 0.75, "t2g",
                       t=-0.26, tp=-0.08, seed=12),
    "two_k4": Workload("two_k4", 2, (4, 4, 2), 4, 2, 8.0, 3.0, 0.5, "rand5",
                       t=0.3, seed=13),
    "d5_k4": Workload("d5_k4", 5, (4, 4, 4), 4, 1, 10.0, 4.0, 0.7, "rand5",
                      t=0.3, seed=14),
    "sc1_k4": Workload("sc1_k4", 1, (4, 4, 4), 8, 2, 10.0, 2.0, 0.0, "sc1", t=0.25, seed=15),
}


def workload(name: str, **overrides) -> Workload:
    wl = WORKLOADS[name]
    if overrides:
        wl = dataclasses.replace(wl, **overrides)
    return wl


@dataclasses.dataclass
class Inputs:
    """Hidden inputs of the hot path (the reference's parameters_module globals)."""
    wl: Workload
    hk: np.ndarray        # (n, n, nkp)
    siw: np.ndarray       # (2*iwmax, n), row iw+iwmax holds nu_iw
    giw: np.ndarray       # (2*iwmax, n)
    dc: np.ndarray        # (2, n); only dc[0, :] enters G
    mu: float
    beta: float
    u: np.ndarray         # (n^2, n^2)
    u_tilde: np.ndarray   # (n^2, n^2)
    n_dmft: np.ndarray    # (n,) complex
    q_data: np.ndarray    # (nqp,) 1-based k indices of the q points
    k_data_eom: np.ndarray  # (nkp_eom,) 1-based k indices of the EOM k points
    v_q: np.ndarray | None = None  # (n^2, n^2, nqp) V(q) in read_vq's compound index, or None

    @property
    def n(self):
        return self.wl.n

    @property
    def iwmax(self):
        return self.siw.shape[0] // 2

    @property
    def iw_data(self):
        i = np.arange(-self.iwmax, self.iwmax)
        return math.pi * (2 * i + 1) / self.beta


def _kgrid(nk):
    nkx, nky, nkz = nk
    ix, iy, iz = np.meshgrid(np.arange(nkx), np.arange(nky), np.arange(nkz), indexing="ij")
    # kz fastest: flat index = kz + ky*nkz + kx*nky*nkz
    kx = 2 * math.pi * ix.reshape(-1) / nkx
    ky = 2 * math.pi * iy.reshape(-1) / nky
    kz = 2 * math.pi * iz.reshape(-1) / nkz
    return kx, ky, kz


def _hk_sc1(wl):
    kx, ky, kz = _kgrid(wl.nk)
    e = -2.0 * wl.t * (np.cos(kx) + np.cos(ky) + np.cos(kz))
    return e.reshape(1, 1, -1).astype(np.complex128)


def _hk_t2g(wl):
    kx, ky, kz = _kgrid(wl.nk)
    t, tp = wl.t, wl.tp
    nkp = kx.size
    h = np.zeros((3, 3, nkp), dtype=np.complex128)
    # orbitals: 0 = xy, 1 = yz, 2 = xz
    h[0, 0] = 2 * t * (np.cos(kx) + np.cos(ky)) + 4 * tp * np.cos(kx) * np.cos(ky)
    h[1, 1] = 2 * t * (np.cos(ky) + np.cos(kz)) + 4 * tp * np.cos(ky) * np.cos(kz)
    h[2, 2] = 2 * t * (np.cos(kx) + np.cos(kz)) + 4 * tp * np.cos(kx) * np.cos(kz)
    # inter-orbital NNN terms make G non-diagonal
    h[1, 2] = h[2, 1] = 4 * tp * np.sin(kx) * np.sin(ky)
    h[0, 2] = h[2, 0] = 4 * tp * np.sin(ky) * np.sin(kz)
    h[0, 1] = h[1, 0] = 4 * tp * np.sin(kx) * np.sin(kz)
    return h


def _hk_rand(wl):
    rng = np.random.default_rng([wl.seed, 7])
    n = wl.n
    kx, ky, kz = _kgrid(wl.nk)
    shells = [(0, 0, 0),
              (1, 0, 0), (0, 1, 0), (0, 0, 1),
              (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1), (0, 1, -1),
              (1, 1, 1), (1, 1, -1), (1, -1, 1), (-1, 1, 1)]
    h = np.zeros((n, n, kx.size), dtype=np.complex128)
    for r in shells:
        a = rng.normal(scale=wl.t, size=(n, n))
        a = 0.5 * (a + a.T)
        if r == (0, 0, 0):
            h += a[:, :, None]
        else:
            ph = np.cos(r[0] * kx + r[1] * ky + r[2] * kz)
            h += 2.0 * a[:, :, None] * ph[None, None, :]
    return h


def kanamori_u(n, U, J):
    """U and U~ in compound index, following create_u (interaction_module.f90:182-291).

    Umat(i,j,k,l): i=j=k=l -> U; i=j, k=l, i!=k -> J (pair hopping);
    i=k, j=l, i!=j -> U-2J; i=l, j=k, i!=j -> J (spin flip).
    u(i1={i,k}, i2={j,l}) = Umat(i,j,k,l), u~(i1, i2) = Umat(i,j,l,k) with
    i1 = i*n + k and i2 = l*n + j (0-based).
    """
    V = U - 2.0 * J
    umat = np.zeros((n, n, n, n))
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    if i == j and k == l:
                        if i == k:
                            umat[i, j, k, l] = U
                            continue
                        umat[i, j, k, l] = J
                        continue
                    if i == k and j == l:
                        umat[i, j, k, l] = V
                        continue
                    if i == l and j == k:
                        umat[i, j, k, l] = J
    u = np.zeros((n * n, n * n), dtype=np.complex128)
    ut = np.zeros_like(u)
    for l in range(n):
        for j in range(n):
            i2 = l * n + j
            for i in range(n):
                for k in range(n):
                    i1 = i * n + k
                    u[i1, i2] = umat[i, j, k, l]
                    ut[i1, i2] = umat[i, j, l, k]
    return u, ut


def k_index(kx, ky, kz, nk):
    """1-based k index (kq_tools.f90:119-125)."""
    return 1 + kz + ky * nk[2] + kx * nk[1] * nk[2]


def q_path(nk):
    """A q path through the k grid like a qdata_from_file list (kq_tools.f90:146-183): Gamma -> X along
    x, X -> M along y, M -> R along z, then R -> Gamma along the diagonal; 1-based k indices."""
    hx, hy, hz = nk[0] // 2, nk[1] // 2, nk[2] // 2
    pts = [(i, 0, 0) for i in range(hx + 1)]
    pts += [(hx, j, 0) for j in range(1, hy + 1)]
    pts += [(hx, hy, k) for k in range(1, hz + 1)]
    m = min(hx, hy, hz)
    pts += [(hx * s // m, hy * s // m, hz * s // m) for s in range(m - 1, -1, -1)]
    return np.array([k_index(a, b, c, nk) for a, b, c in pts], dtype=np.int32)


def synthetic_vq(wl, q_data):
    """V(q)(i1 = {i,k}, i2 = {l,j}) = vq(i,j,k,l) (read_vq, interaction_module.f90:106-119): a
    density-density nearest-neighbour interaction V_ij(q) = V (cos qx + cos qy + cos qz)/3, full weight
    within an orbital and 0.6 between orbitals, plus a small seeded exchange part vq(i,j,j,i) = 0.1 V_ij(q)
    for i != j; real and symmetric under q -> -q."""
    n = wl.n
    q = np.asarray(q_data) - 1
    nkx, nky, nkz = wl.nk
    qz, qy, qx = q % nkz, (q // nkz) % nky, q // (nky * nkz)
    c = (np.cos(2 * math.pi * qx / nkx) + np.cos(2 * math.pi * qy / nky) + np.cos(2 * math.pi * qz / nkz)) / 3.0
    rng = np.random.default_rng([wl.seed, 9])
    w = 0.6 + 0.05 * rng.normal(size=(n, n))
    w = 0.5 * (w + w.T)
    np.fill_diagonal(w, 1.0)
    v = np.zeros((n * n, n * n, q.size), dtype=np.complex128)
    for i in range(n):
        for j in range(n):
            v[i * n + i, j * n + j, :] = wl.vq * w[i, j] * c          # vq(i,j,i,j)
            if i != j:
                v[i * n + j, i * n + j, :] = 0.1 * wl.vq * w[i, j] * c  # vq(i,j,j,i): i1 = {i,j}, i2 = {i,j}
    return v


def generate_q_vol(nk, nq):
    """q list on a sub-grid of the k grid (kq_tools.f90:42-57), 1-based."""
    out = []
    for i in range(nq[0]):
        for j in range(nq[1]):
            for k in range(nq[2]):
                out.append(k_index(i * nk[0] // nq[0], j * nk[1] // nq[1], k * nk[2] // nq[2], nk))
    return np.array(out, dtype=np.int32)


def k_minus_q(ik, iq, nk):
    """1-based index of k - q with fold-back by integer modulo (kq_tools.f90:96-108)."""
    ik = np.asarray(ik) - 1
    iq = np.asarray(iq) - 1
    nkx, nky, nkz = nk
    iz, iy, ix = ik % nkz, (ik // nkz) % nky, ik // (nky * nkz)
    lz, ly, lx = iq % nkz, (iq // nkz) % nky, iq // (nky * nkz)
    return (1 + (nkz + iz - lz) % nkz + ((nky + iy - ly) % nky) * nkz
            + ((nkx + ix - lx) % nkx) * nky * nkz)


def _hubbard_i(wl):
    """Parameters (h, s, e) of the orbital-diagonal Hubbard-I-like self-energy h + s^2 / (i nu - e); each is
    proportional to U."""
    n = wl.n
    rng = np.random.default_rng([wl.seed, 1])
    if n == 1:
        return np.array([wl.U / 2]), np.array([wl.U / 2]), np.zeros(1)
    h = np.full(n, wl.U / 2) + 0.02 * wl.U * rng.normal(size=n)
    s = wl.U / 2 * (1.0 + 0.05 * rng.normal(size=n))
    e = 0.05 * wl.U * rng.normal(size=n)
    return h, s, e


def _g_blocks(hk, siw, dc, mu, nu, iwmax, block=64):
    """G(k, nu) for nu >= 0 in blocks of frequencies: yields (ws, G[w, k, a, b])."""
    n = hk.shape[0]
    hkt = np.ascontiguousarray(hk.transpose(2, 0, 1))  # (nkp, n, n)
    eye = np.eye(n)
    for i0 in range(0, iwmax, block):
        ws = np.arange(i0, min(iwmax, i0 + block))
        a = -hkt[None, :, :, :] + np.zeros((ws.size, 1, 1, 1))
        diag = (1j * nu[ws + iwmax][:, None] + mu - dc[0][None, :] - siw[ws + iwmax])  # (nw, n)
        a = a + diag[:, None, :, None] * eye[None, None, :, :]
        yield ws, np.linalg.inv(a)


def _hk(wl):
    if wl.model == "sc1":
        return _hk_sc1(wl)
    if wl.model == "t2g":
        return _hk_t2g(wl)
    return _hk_rand(wl)


PARAMS = ("t", "tp", "U")


@dataclasses.dataclass
class Tangents:
    """Derivatives of the hidden inputs with respect to the model parameters PARAMS = (t, t', U), the
    directions of the GPU path's tangent-linear mode: dhk(n, n, nkp, 3), dsiw and dgiw (2*iwmax, n, 3),
    dmu (3,), du and du_tilde (n^2, n^2, 3).  dc and the chi_loc noise do not depend on them."""
    dhk: np.ndarray
    dsiw: np.ndarray
    dgiw: np.ndarray
    dmu: np.ndarray
    du: np.ndarray
    du_tilde: np.ndarray


def tangents(inp: Inputs) -> Tangents:
    """d(input)/d(t, t', U) of build_inputs: H(k) is linear in t and t' (t2g, sc1) or proportional to t
    (random hopping); h, s, e of the self-energy and mu = U/2 are proportional to U; U and U~ are linear in
    U at fixed J; dG = G (dH + dSigma - dmu) G gives dgiw = 1/nkp sum_k diag dG."""
    wl = inp.wl
    n, iwmax = wl.n, inp.iwmax
    nkp = wl.nkp
    dhk = np.zeros((n, n, nkp, 3), dtype=np.complex128)
    if wl.model == "rand5":
        dhk[..., 0] = inp.hk / wl.t if wl.t != 0 else 0.0
    else:
        dhk[..., 0] = _hk(dataclasses.replace(wl, t=1.0, tp=0.0))
        if wl.model == "t2g":
            dhk[..., 1] = _hk(dataclasses.replace(wl, t=0.0, tp=1.0))
    h, s, e = _hubbard_i(wl)
    nu = inp.iw_data
    den = 1j * nu[:, None] - e[None, :]
    dsiw = np.zeros((2 * iwmax, n, 3), dtype=np.complex128)
    dsiw[..., 2] = (h[None, :] + 2 * s[None, :] ** 2 / den + s[None, :] ** 2 * e[None, :] / den ** 2) / wl.U
    dmu = np.array([0.0, 0.0, 0.5])
    dgiw = np.zeros((2 * iwmax, n, 3), dtype=np.complex128)
    hkt = np.ascontiguousarray(dhk.transpose(2, 3, 0, 1))  # (nkp, 3, n, n)
    for ws, g in _g_blocks(inp.hk, inp.siw, inp.dc, inp.mu, nu, iwmax):
        for d in range(3):
            x = np.broadcast_to(hkt[None, :, d], g.shape).copy()
            x[..., np.arange(n), np.arange(n)] += (dsiw[ws + iwmax, :, d] - dmu[d])[:, None, :]
            dg = g @ x @ g
            dgiw[ws + iwmax, :, d] = np.einsum("wkaa->wa", dg) / nkp
    for w in range(iwmax):
        dgiw[-w - 1 + iwmax] = np.conj(dgiw[w + iwmax])
    du = np.zeros((n * n, n * n, 3), dtype=np.complex128)
    dut = np.zeros_like(du)
    du[..., 2], dut[..., 2] = kanamori_u(n, 1.0, 0.0)
    return Tangents(np.asfortranarray(dhk), np.asfortranarray(dsiw), np.asfortranarray(dgiw), dmu,
                    np.asfortranarray(du), np.asfortranarray(dut))


def dchi_loc(inp: Inputs, tan: Tangents, iwb: int, box: int | None = None) -> np.ndarray:
    """d chi_loc_r(omega) / d(t, t', U), (N, N, 3), the same for both channels: the local bubble's
    -beta d(g_i(nu) g_j(nu - omega)) on the diagonal (the noise is fixed)."""
    wl = inp.wl
    n = wl.n
    iwf = wl.iwfmax_small if box is None else box
    nf = 2 * iwf
    N = n * n * nf
    off = inp.iwmax
    out = np.zeros((N, N, 3), dtype=np.complex128, order="F")
    for d in range(nf):
        w = d - iwf
        gi, gj = inp.giw[w + off], inp.giw[w - iwb + off]
        for p in range(3):
            dgi, dgj = tan.dgiw[w + off, :, p], tan.dgiw[w - iwb + off, :, p]
            v = -inp.beta * (dgi[:, None] * gj[None, :] + gi[:, None] * dgj[None, :])
            idx = d * n * n + np.arange(n * n)
            out[idx, idx, p] = v.reshape(-1)
    return out


def build_inputs(wl: Workload) -> Inputs:
    n = wl.n
    beta = wl.beta
    iwmax = wl.n_iwmax
    if iwmax < wl.iwfmax_small + wl.iwbmax_small:
        raise ValueError("iwmax must be >= iwfmax_small + iwbmax_small (config_module.f90:600-607)")
    hk = _hk(wl)
    assert hk.shape[0] == n
    h, s, e = _hubbard_i(wl)
    mu = float(wl.U / 2)
    iw = np.arange(-iwmax, iwmax)
    nu = math.pi * (2 * iw + 1) / beta
    siw = h[None, :] + s[None, :] ** 2 / (1j * nu[:, None] - e[None, :])
    dc = np.zeros((2, n), dtype=np.complex128)
    # g(nu) = 1/nkp sum_k diag G(k, nu) for nu >= 0, conjugate for nu < 0 (get_giw)
    giw = np.zeros((2 * iwmax, n), dtype=np.complex128)
    for ws, g in _g_blocks(hk, siw, dc, mu, nu, iwmax):
        giw[ws + iwmax] = np.einsum("wkaa->wa", g) / hk.shape[2]
    for w in range(iwmax):
        giw[-w - 1 + iwmax] = np.conj(giw[w + iwmax])
    # occupations (get_ndmft): n = 2 Re sum_{nu>=0} g / beta + 1/2
    n_dmft = (2.0 * np.real(giw[iwmax:].sum(axis=0)) / beta + 0.5).astype(np.complex128)
    u, ut = kanamori_u(n, wl.U, wl.J)
    q_data = q_path(wl.nk) if wl.q_list == "path" else generate_q_vol(wl.nk, wl.qgrid)
    if wl.k_eom == "all":
        k_eom = np.arange(1, wl.nkp + 1, dtype=np.int32)
    else:
        k_eom = np.arange(wl.nkp, 0, -3, dtype=np.int32)
    v_q = np.asfortranarray(synthetic_vq(wl, q_data)) if wl.vq != 0.0 else None
    return Inputs(wl, np.asfortranarray(hk), np.asfortranarray(siw), np.asfortranarray(giw),
                  np.asfortranarray(dc), mu, beta, np.asfortranarray(u), np.asfortranarray(ut),
                  n_dmft, q_data, k_eom, v_q)


def chi_loc(inp: Inputs, iwb: int, channel: str) -> np.ndarray:
    """Local generalized susceptibility chi_loc_r(omega) in the compound index that
    `read_vertex` produces (hdf5_module.f90:708-770):
    row i1 = nu1_idx*n^2 + i*n + j, column i2 = nu2_idx*n^2 + l*n + k (0-based).
    Local bubble -beta g_i(nu) g_j(nu-omega) on the diagonal plus seeded Hermitian noise.
    """
    wl = inp.wl
    n, nf = wl.n, wl.nf
    N = n * n * nf
    off = inp.iwmax
    rch = {"dens": 0, "magn": 1}[channel]
    rng = np.random.default_rng([wl.seed, 100 + rch, 1000 + iwb])
    x = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    a = wl.noise * 0.5 * (x + x.conj().T)
    diag = np.empty(N, dtype=np.complex128)
    for d in range(nf):
        iwf = d - wl.iwfmax_small
        gi = inp.giw[iwf + off]
        gj = inp.giw[iwf - iwb + off]
        diag[d * n * n:(d + 1) * n * n] = (-inp.beta * gi[:, None] * gj[None, :]).reshape(-1)
    a[np.arange(N), np.arange(N)] += diag
    return np.asfortranarray(a)
