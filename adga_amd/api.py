"""Host-side mirror of the reference's hot-path interface over the C ABI.

`Context` owns one device context (adga_ctx_create).  Its methods follow the
reference's loop body (src/main.F90:486-799): `set_local_omega` is the local
update done when omega changes (506-625), `run_points` the per-point work for a
batch of q points (635-781), `get_sigma` the accumulated self-energies (841-855).
`run_loop` drives a whole (omega, q) point list the way main.F90 does and
returns the same quantities the reference accumulates.  `get_giw`, `get_nfock`,
`get_sigma_g_loc` and `get_ndga` are the one-particle routines around the loop
(src/one_particle_quant_module.f90).  `inverse_matrix` mirrors lapack_module's
inverse_matrix (src/lapack_module.f90:40-70).

Errors follow the reference's fatal-error behaviour (mpi_stop): a non-zero
status raises `AdgaError` carrying the status code and the library's message.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import ADGA_OK, ADGA_PREC_FP64, ADGA_PREC_REFINED, lib

__all__ = ["AdgaError", "Context", "Lambda", "run_loop", "run_loop_tangent", "inverse_matrix", "PRECISION"]

PRECISION = {"fp64": ADGA_PREC_FP64, "refined": ADGA_PREC_REFINED}


class AdgaError(RuntimeError):
    def __init__(self, code, msg, status=None):
        super().__init__(f"adga status {code}: {msg}")
        self.code = code
        self.status = status


def _zf(a):
    """Fortran-ordered complex128 copy/view."""
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


def _dptr(a):
    return a.ctypes.data_as(_lib._dp)


class Context:
    def __init__(self, inp, batch=None, refine_tol=None, refine_maxit=None, bubble=None, eom=None, lu_streams=None,
                 do_eom=None, fallback=None, bubble_range=None):
        wl = inp.wl
        self.n = wl.n
        self.n2 = wl.n * wl.n
        self.nf = wl.nf
        self.iwfmax_small = wl.iwfmax_small
        self.iwbmax_small = wl.iwbmax_small
        self.nkp_eom = len(inp.k_data_eom)
        self.nqp = len(inp.q_data)
        self.iwmax = int(inp.iwmax)
        self.nkp = int(wl.nk[0] * wl.nk[1] * wl.nk[2])
        d = _lib.adga_dims(wl.n, wl.nk[0], wl.nk[1], wl.nk[2], inp.iwmax, wl.iwfmax_small, wl.iwbmax_small,
                           self.nqp, self.nkp_eom, float(inp.beta), float(inp.mu))
        self._keep = [_zf(inp.hk), _zf(inp.siw), _zf(inp.giw), _zf(inp.dc), _zf(inp.u), _zf(inp.u_tilde),
                      _zf(inp.n_dmft)]
        q = np.ascontiguousarray(inp.q_data, dtype=np.int32)
        k = np.ascontiguousarray(inp.k_data_eom, dtype=np.int32)
        h = C.c_void_p()
        rc = lib.adga_ctx_create(C.byref(d), q.ctypes.data_as(_lib._ip), k.ctypes.data_as(_lib._ip),
                                 *[_dptr(a) for a in self._keep], C.byref(h))
        if rc != ADGA_OK:
            raise AdgaError(rc, "adga_ctx_create failed")
        self.h = h
        self._keep = None
        if getattr(inp, "v_q", None) is not None:
            self.set_vq(inp.v_q)
        if do_eom is None:
            do_eom = getattr(wl, "do_eom", True)  # the reference runs a q path without the EOM
        if bubble_range is None:
            bubble_range = getattr(wl, "bubble_range", 0) or None  # the debug option "Bubble" (+ the tail)
        for key, v in (("batch", batch), ("refine_tol", refine_tol), ("refine_maxit", refine_maxit),
                       ("bubble", bubble), ("eom", eom), ("lu_streams", lu_streams), ("do_eom", int(bool(do_eom))),
                       ("fallback", fallback), ("bubble_range", bubble_range)):
            if v is not None:
                self._check(lib.adga_set_option(self.h, key.encode(), float(v)))

    # -----------------------------------------------------------------
    def _err(self):
        buf = C.create_string_buffer(512)
        lib.adga_last_error(self.h, buf, 512)
        return buf.value.decode(errors="replace")

    def _check(self, rc, status=None):
        if rc != ADGA_OK:
            raise AdgaError(rc, self._err(), status)

    def close(self):
        if getattr(self, "h", None):
            lib.adga_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -----------------------------------------------------------------
    def set_vq(self, v_q):
        """V(q) for every q of the list, (n2, n2, nqp) as read_vq returns it per point
        (src/interaction_module.f90:39-123); None removes it (do_vq = .false.)."""
        if v_q is None:
            self._check(lib.adga_set_vq(self.h, None))
            return
        v = _zf(v_q)
        if v.shape != (self.n2, self.n2, self.nqp):
            raise ValueError(f"v_q must be ({self.n2}, {self.n2}, {self.nqp})")
        self._check(lib.adga_set_vq(self.h, _dptr(v)))

    def set_local_omega(self, iwb, chi_loc_dens, chi_loc_magn, box=None):
        """Local update at omega = iwb; `box` (default iwfmax_small): this omega's fermionic box
        nu in [-box, box), with chi_loc of size n^2 * 2 * box."""
        n2 = self.n2
        box = self.iwfmax_small if box is None else int(box)
        xd, xm = _zf(chi_loc_dens), _zf(chi_loc_magn)
        N = n2 * 2 * box
        if xd.shape != (N, N) or xm.shape != (N, N):
            raise ValueError(f"chi_loc must be ({N}, {N})")
        od = np.zeros((n2, n2), dtype=np.complex128, order="F")
        om = np.zeros_like(od)
        bl = np.zeros_like(od)
        self._check(lib.adga_set_local_omega_box(self.h, int(iwb), box, _dptr(xd), _dptr(xm), _dptr(od), _dptr(om),
                                                 _dptr(bl)))
        return od, om, bl

    def run_points(self, iq, precision="fp64", raise_on_error=True):
        """iq: 1-based positions in q_data.  Returns chi_d, chi_m, bubble_nl of shape (n2, n2, npts)
        and the per-point status."""
        iq = np.ascontiguousarray(iq, dtype=np.int32)
        npts = iq.size
        n2 = self.n2
        cd = np.zeros((n2, n2, npts), dtype=np.complex128, order="F")
        cm = np.zeros_like(cd)
        bn = np.zeros_like(cd)
        st = np.zeros(npts, dtype=np.int32)
        rc = lib.adga_run_points(self.h, npts, iq.ctypes.data_as(_lib._ip), PRECISION[precision], _dptr(cd),
                                 _dptr(cm), _dptr(bn), st.ctypes.data_as(_lib._ip))
        if raise_on_error:
            self._check(rc, st)
        return {"chi_d": cd, "chi_m": cm, "bubble_nl": bn, "status": st, "rc": rc}

    def get_sigma(self):
        n = self.n
        snl = np.zeros((n, n, self.nf, self.nkp_eom), dtype=np.complex128, order="F")
        sdm = np.zeros((n, n, self.nf), dtype=np.complex128, order="F")
        shf = np.zeros((n, n, self.nkp_eom), dtype=np.complex128, order="F")
        self._check(lib.adga_get_sigma(self.h, _dptr(snl), _dptr(sdm), _dptr(shf)))
        return snl, sdm, shf

    # ----------------------------------------------------------------- one-particle routines
    def get_giw(self):
        """giw(-iwmax:iwmax-1, ndim) like get_giw (src/one_particle_quant_module.f90:40-79)."""
        out = np.zeros((2 * self.iwmax, self.n), dtype=np.complex128, order="F")
        self._check(lib.adga_get_giw(self.h, _dptr(out)))
        return out

    def get_nfock(self):
        """n_fock(nkp, ndim, ndim) like get_nfock (src/one_particle_quant_module.f90:337-361)."""
        out = np.zeros((self.nkp, self.n, self.n), dtype=np.complex128, order="F")
        self._check(lib.adga_get_nfock(self.h, _dptr(out)))
        return out

    def get_sigma_g_loc(self, sigma_sum=None):
        """(sigma_loc, gloc) like get_sigma_g_loc (src/one_particle_quant_module.f90:124-180);
        sigma_sum(ndim, ndim, 2*iwfmax_small, nkp) after add_siw_dmft, or None for the context's own
        (its accumulated Sigma_nl plus siw, formed on the device)."""
        if sigma_sum is None:
            ptr = None
        else:
            ss = _zf(sigma_sum)
            if ss.shape != (self.n, self.n, self.nf, self.nkp):
                raise ValueError(f"sigma_sum must be ({self.n}, {self.n}, {self.nf}, {self.nkp})")
            ptr = _dptr(ss)
        sloc = np.zeros((self.n, self.n, self.nf), dtype=np.complex128, order="F")
        gloc = np.zeros((2 * self.iwmax, self.n, self.n), dtype=np.complex128, order="F")
        self._check(lib.adga_get_sigma_g_loc(self.h, ptr, _dptr(sloc), _dptr(gloc)))
        return sloc, gloc

    def get_ndga(self, sigma_sum=None):
        """(n_dga_k, n_dga) like get_ndga (src/one_particle_quant_module.f90:363-393);
        sigma_sum(ndim, ndim, 2*iwfmax_small, nkp_eom) after add_siw_dmft, or None for the context's own."""
        if sigma_sum is None:
            ptr = None
        else:
            ss = _zf(sigma_sum)
            if ss.shape != (self.n, self.n, self.nf, self.nkp_eom):
                raise ValueError(f"sigma_sum must be ({self.n}, {self.n}, {self.nf}, {self.nkp_eom})")
            ptr = _dptr(ss)
        ndk = np.zeros((self.nkp_eom, self.n, self.n), dtype=np.complex128, order="F")
        nd = np.zeros(self.n, dtype=np.complex128)
        self._check(lib.adga_get_ndga(self.h, ptr, _dptr(ndk), _dptr(nd)))
        return ndk, nd

    # ----------------------------------------------------------------- derivatives (tangent-linear mode)
    def set_tangents(self, tan):
        """Directions of the inputs to differentiate along: a synth.Tangents (or any object with dhk, dsiw,
        dgiw, dmu, du, du_tilde in those shapes, the last axis the direction)."""
        self.ndir = tan.dhk.shape[-1]
        arrs = [_zf(tan.dhk), _zf(tan.dsiw), _zf(tan.dgiw)]
        dmu = np.ascontiguousarray(tan.dmu, dtype=np.float64)
        du, dut = _zf(tan.du), _zf(tan.du_tilde)
        self._check(lib.adga_set_tangents(self.h, self.ndir, *[_dptr(a) for a in arrs], _dptr(dmu), _dptr(du),
                                          _dptr(dut)))

    def set_local_omega_tangent(self, dchi_loc_dens=None, dchi_loc_magn=None):
        """After set_local_omega: derivatives of its local data given those of chi_loc (N, N, ndir) or None;
        returns the derivatives of (chi_loc_d, chi_loc_m, bubble_loc), each (n2, n2, ndir)."""
        n2, nd = self.n2, self.ndir
        xd = None if dchi_loc_dens is None else _zf(dchi_loc_dens)
        xm = None if dchi_loc_magn is None else _zf(dchi_loc_magn)
        od = np.zeros((n2, n2, nd), dtype=np.complex128, order="F")
        om, bl = np.zeros_like(od), np.zeros_like(od)
        self._check(lib.adga_set_local_omega_tangent(self.h, None if xd is None else _dptr(xd),
                                                     None if xm is None else _dptr(xm), _dptr(od), _dptr(om),
                                                     _dptr(bl)))
        return od, om, bl

    def run_points_tangent(self, iq):
        """run_points in complex128 with the derivatives dchi_d, dchi_m, dbubble_nl (n2, n2, npts, ndir)."""
        iq = np.ascontiguousarray(iq, dtype=np.int32)
        npts, n2, nd = iq.size, self.n2, self.ndir
        out = {k: np.zeros((n2, n2, npts), dtype=np.complex128, order="F") for k in ("chi_d", "chi_m", "bubble_nl")}
        for k in ("dchi_d", "dchi_m", "dbubble_nl"):
            out[k] = np.zeros((n2, n2, npts, nd), dtype=np.complex128, order="F")
        st = np.zeros(npts, dtype=np.int32)
        rc = lib.adga_run_points_tangent(self.h, npts, iq.ctypes.data_as(_lib._ip), _dptr(out["chi_d"]),
                                         _dptr(out["chi_m"]), _dptr(out["bubble_nl"]), _dptr(out["dchi_d"]),
                                         _dptr(out["dchi_m"]), _dptr(out["dbubble_nl"]), st.ctypes.data_as(_lib._ip))
        self._check(rc, st)
        out["status"] = st
        return out

    def get_sigma_tangent(self):
        n = self.n
        d = np.zeros((n, n, self.nf, self.nkp_eom, self.ndir), dtype=np.complex128, order="F")
        self._check(lib.adga_get_sigma_tangent(self.h, _dptr(d)))
        return d

    def reset_sigma(self):
        self._check(lib.adga_reset_sigma(self.h))

    def stats(self):
        out = np.zeros(16)
        self._check(lib.adga_get_stats(self.h, _dptr(out), 16))
        keys = ["refine_iters", "refine_resid", "t_local", "t_chi0", "t_gammas", "t_inversion", "t_eta_rest",
                "t_eom", "t_chi", "t_lu_kernel", "n_lu_launches", "t_gemm_dominant", "flops_gemm_dominant",
                "n_gemm_dominant", "n_fp64_resolved", "n_failed_points"]
        return dict(zip(keys, out.tolist()))


def run_loop(ctx: Context, points, chi_loc_fn, precision="fp64", box_fn=None):
    """Drive `points` = [(iwb, iq0), ...] (iq0 0-based, omega-major like main.F90:381-389) through the
    device path.  Consecutive points with the same omega form one batch.  Returns the quantities the
    reference accumulates: per-point chi_d, chi_m, bubble_nl, and sigma_nl, sigma_dmft, chi_loc_d/m and
    bubble_loc per omega (added when the omega block starts at iq == 1, main.F90:519-577).  `box_fn(iwb)`
    (optional) gives each omega's fermionic box; chi_loc_fn then returns chi_loc of that box."""
    n2 = ctx.n2
    npts = len(points)
    out = {"chi_d": np.zeros((n2, n2, npts), dtype=np.complex128, order="F"),
           "chi_m": np.zeros((n2, n2, npts), dtype=np.complex128, order="F"),
           "bubble_nl": np.zeros((n2, n2, npts), dtype=np.complex128, order="F"),
           "status": np.zeros(npts, dtype=np.int32),
           "chi_loc_d": {}, "chi_loc_m": {}, "bubble_loc": {}}
    p = 0
    while p < npts:
        iwb = points[p][0]
        e = p
        while e < npts and points[e][0] == iwb:
            e += 1
        box = None if box_fn is None else box_fn(iwb)
        ld, lm, bl = ctx.set_local_omega(iwb, chi_loc_fn(iwb, "dens"), chi_loc_fn(iwb, "magn"), box)
        if points[p][1] == 0:
            out["chi_loc_d"][iwb] = ld
            out["chi_loc_m"][iwb] = lm
            out["bubble_loc"][iwb] = bl
        iq = np.array([q + 1 for (_, q) in points[p:e]], dtype=np.int32)
        r = ctx.run_points(iq, precision)
        out["chi_d"][:, :, p:e] = r["chi_d"]
        out["chi_m"][:, :, p:e] = r["chi_m"]
        out["bubble_nl"][:, :, p:e] = r["bubble_nl"]
        out["status"][p:e] = r["status"]
        p = e
    out["sigma_nl"], out["sigma_dmft"], out["sigma_hf"] = ctx.get_sigma()
    return out


def run_loop_tangent(ctx: Context, points, chi_loc_fn, dchi_loc_fn, box_fn=None):
    """run_loop in complex128 with the derivatives along the context's tangent directions
    (Context.set_tangents): `dchi_loc_fn(iwb, channel)` gives d chi_loc (N, N, ndir) of a channel.  Returns
    run_loop's quantities and d<name> for each of them (last axis: the direction)."""
    n2, nd = ctx.n2, ctx.ndir
    npts = len(points)
    out = {k: np.zeros((n2, n2, npts), dtype=np.complex128, order="F") for k in ("chi_d", "chi_m", "bubble_nl")}
    for k in ("dchi_d", "dchi_m", "dbubble_nl"):
        out[k] = np.zeros((n2, n2, npts, nd), dtype=np.complex128, order="F")
    out["status"] = np.zeros(npts, dtype=np.int32)
    for k in ("chi_loc_d", "chi_loc_m", "bubble_loc", "dchi_loc_d", "dchi_loc_m", "dbubble_loc"):
        out[k] = {}
    p = 0
    while p < npts:
        iwb = points[p][0]
        e = p
        while e < npts and points[e][0] == iwb:
            e += 1
        box = None if box_fn is None else box_fn(iwb)
        ld, lm, bl = ctx.set_local_omega(iwb, chi_loc_fn(iwb, "dens"), chi_loc_fn(iwb, "magn"), box)
        dld, dlm, dbl = ctx.set_local_omega_tangent(dchi_loc_fn(iwb, "dens"), dchi_loc_fn(iwb, "magn"))
        if points[p][1] == 0:
            for k, v in (("chi_loc_d", ld), ("chi_loc_m", lm), ("bubble_loc", bl), ("dchi_loc_d", dld),
                         ("dchi_loc_m", dlm), ("dbubble_loc", dbl)):
                out[k][iwb] = v
        r = ctx.run_points_tangent(np.array([q + 1 for (_, q) in points[p:e]], dtype=np.int32))
        for k in ("chi_d", "chi_m", "bubble_nl", "dchi_d", "dchi_m", "dbubble_nl"):
            out[k][:, :, p:e] = r[k]
        out["status"][p:e] = r["status"]
        p = e
    out["sigma_nl"], out["sigma_dmft"], out["sigma_hf"] = ctx.get_sigma()
    out["dsigma_nl"] = ctx.get_sigma_tangent()
    return out


def inverse_matrix(m):
    """In-place-style inverse like inverse_matrix_z (src/lapack_module.f90:40-70): returns (m^-1, ierr).
    Raises AdgaError for a singular matrix (the reference stops with "ERROR in ZGETRF")."""
    a = np.array(m, dtype=np.complex128, order="F", copy=True)
    if a.ndim != 2 or a.shape[0] != a.shape[1]:
        raise ValueError("square matrix expected")
    ierr = C.c_int(0)
    rc = lib.adga_inverse_matrix_z(a.shape[0], _dptr(a), C.byref(ierr))
    if rc != ADGA_OK:
        raise AdgaError(rc, f"ERROR in ZGETRF (info={ierr.value})" if rc == _lib.ADGA_E_SINGULAR else "inverse failed")
    return a, ierr.value


class Lambda:
    """lambda correction of the final susceptibility on the device (adga_lambda_*; no reference counterpart).
    chi_qw (n2, n2, nqp*(2*iwbmax+1)) over the omega-major (omega, q) list and chi_loc (n2, n2, 2*iwbmax+1),
    both the final ones (with the bubbles); the physical susceptibility is sum_{i,j} chi[(i,i), (j,j)]."""

    def __init__(self, chi_qw, chi_loc, beta, nqp):
        n2 = chi_qw.shape[0]
        n = int(round(n2 ** 0.5))
        nw = chi_loc.shape[2]
        if chi_qw.shape != (n2, n2, nqp * nw) or n * n != n2:
            raise ValueError("chi_qw must be (n^2, n^2, nqp * n_omega) with chi_loc (n^2, n^2, n_omega)")
        self.nqp, self.nw, self.beta = nqp, nw, float(beta)
        self._cq = _zf(chi_qw)
        self._cl = _zf(chi_loc)
        h = C.c_void_p()
        rc = lib.adga_lambda_create(n, nqp, (nw - 1) // 2, self.beta, _dptr(self._cq), _dptr(self._cl), C.byref(h))
        if rc != ADGA_OK:
            raise AdgaError(rc, "adga_lambda_create failed")
        self.h = h
        self._cq = self._cl = None

    def _check(self, rc):
        if rc != ADGA_OK:
            buf = C.create_string_buffer(512)
            lib.adga_lambda_last_error(self.h, buf, 512)
            raise AdgaError(rc, buf.value.decode(errors="replace"))

    def sums(self, lambdas):
        """(S(lambda), dS/dlambda) for each lambda, complex arrays."""
        lam = np.ascontiguousarray(np.atleast_1d(lambdas), dtype=np.float64)
        s = np.zeros(lam.size, dtype=np.complex128)
        ds = np.zeros_like(s)
        self._check(lib.adga_lambda_sums(self.h, lam.size, lam.ctypes.data_as(_lib._dp), _dptr(s), _dptr(ds)))
        return s, ds

    def target(self):
        t = C.c_double()
        self._check(lib.adga_lambda_target(self.h, C.byref(t)))
        return t.value

    def solve(self, target=None, tol=1e-14, maxit=100):
        """(lambda, iterations) with Re S(lambda) = target (the local sum rule by default)."""
        t = self.target() if target is None else float(target)
        lam = C.c_double()
        it = C.c_int()
        self._check(lib.adga_lambda_solve(self.h, t, tol, maxit, C.byref(lam), C.byref(it)))
        return lam.value, it.value

    def chi(self, lam):
        out = np.zeros(self.nqp * self.nw, dtype=np.complex128)
        self._check(lib.adga_lambda_chi(self.h, float(lam), _dptr(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib.adga_lambda_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
