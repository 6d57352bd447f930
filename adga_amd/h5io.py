"""ctypes binding of libadga_h5.so (include/adga_h5.h): the reference's HDF5 file formats.

`H5File` mirrors the reference's readers (src/hdf5_module.f90: read_siw, read_giw, read_dc, read_mu,
read_beta, read_hk_w2dyn, read_vertex; src/interaction_module.f90: read_vq) and writers (init_h5_output,
output_chi_qw_*_h5, output_chi_qpath_*_h5, output_chi_loc_*_h5, output_eom_h5, output_occ_h5).  Arrays are
complex128 in the reference's Fortran shapes.  Errors raise `H5Error`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libadga_h5.so")

EXPORTS = ["adga_h5_open", "adga_h5_create", "adga_h5_close", "adga_h5_last_error", "adga_h5_read_siw",
           "adga_h5_read_giw", "adga_h5_read_dc", "adga_h5_read_mu", "adga_h5_read_beta", "adga_h5_read_hk",
           "adga_h5_read_vertex", "adga_h5_write_vertex", "adga_h5_read_vq", "adga_h5_init_output", "adga_h5_write_chi_qw",
           "adga_h5_write_chi_qpath", "adga_h5_write_chi_loc", "adga_h5_write_eom", "adga_h5_write_occ",
           "adga_h5_dims", "adga_h5_read_complex", "adga_h5_read_real", "adga_h5_write_complex", "adga_h5_write_real",
           "adga_h5_write_attr_real"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Atoms(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nineq", C.c_int), ("ndims", _ip)]


class Vertex(C.Structure):
    _fields_ = [("atoms", Atoms), ("iwfmax", C.c_int), ("iwbmax", C.c_int), ("iwfmax_small", C.c_int),
                ("iwmax", C.c_int), ("vertex_type", C.c_int), ("beta", C.c_double), ("giw", _dp)]


class Run(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("ndim", "iwmax", "iwfmax", "iwbmax", "iwfmax_small", "iwbmax_small", "nkp",
                                       "nkpx", "nkpy", "nkpz", "nqp", "nqpx", "nqpy", "nqpz", "nkp_eom", "do_chi",
                                       "do_eom", "nonlocal", "gamma_group", "q_path_susc", "k_path_eom")] + \
               [("mu", C.c_double), ("beta", C.c_double), ("dc", _dp), ("giw", _dp), ("siw", _dp), ("hk", _dp),
                ("n_dmft", _dp), ("n_fock", _dp), ("k_data", _dp), ("q_data", _ip), ("k_data_eom", _ip)]


def _load():
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is not built: run `python -m adga_amd.build` (needs the HDF5 C library)")
    lib = C.CDLL(LIBPATH)
    vp = C.c_void_p
    sig = {
        "adga_h5_open": [C.c_char_p, C.c_int, C.POINTER(vp)],
        "adga_h5_create": [C.c_char_p, C.POINTER(vp)],
        "adga_h5_close": [vp],
        "adga_h5_last_error": [vp, C.c_char_p, C.c_int],
        "adga_h5_read_siw": [vp, C.c_char_p, C.POINTER(Atoms), C.c_int, C.c_int, _dp],
        "adga_h5_read_giw": [vp, C.c_char_p, C.POINTER(Atoms), C.c_int, C.c_int, _dp],
        "adga_h5_read_dc": [vp, C.c_char_p, C.POINTER(Atoms), _dp],
        "adga_h5_read_mu": [vp, C.c_char_p, _dp],
        "adga_h5_read_beta": [vp, _dp],
        "adga_h5_read_hk": [vp, _ip, _ip, _dp],
        "adga_h5_read_vertex": [vp, C.c_int, C.POINTER(Vertex), _dp, _dp],
        "adga_h5_write_vertex": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp],
        "adga_h5_read_vq": [vp, C.c_int, C.c_int, C.c_int, _ip, _dp],
        "adga_h5_init_output": [vp, C.POINTER(Run)],
        "adga_h5_write_chi_qw": [vp, C.c_char_p, C.c_int, C.c_int, C.c_int, _ip, _dp, C.c_int],
        "adga_h5_write_chi_qpath": [vp, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp],
        "adga_h5_write_chi_loc": [vp, C.c_char_p, C.c_int, C.c_int, C.c_int, _dp, C.c_int],
        "adga_h5_write_eom": [vp, C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp, C.c_int],
        "adga_h5_write_occ": [vp, C.c_int, _ip, _dp, _dp],
        "adga_h5_dims": [vp, C.c_char_p, _ip, C.POINTER(C.c_longlong)],
        "adga_h5_read_complex": [vp, C.c_char_p, _dp],
        "adga_h5_read_real": [vp, C.c_char_p, _dp],
        "adga_h5_write_complex": [vp, C.c_char_p, C.c_int, C.POINTER(C.c_longlong), _dp],
        "adga_h5_write_real": [vp, C.c_char_p, C.c_int, C.POINTER(C.c_longlong), _dp],
        "adga_h5_write_attr_real": [vp, C.c_char_p, C.c_char_p, C.c_double],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    return lib


lib = _load()


class H5Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"adga_h5 status {code}: {msg}")
        self.code = code


def _z(a):
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _atoms(ndim, ndims):
    nd = np.asfortranarray(np.asarray(ndims, dtype=np.int32).reshape(-1, 2))
    return Atoms(ndim, nd.shape[0], nd.ctypes.data_as(_ip)), nd


class H5File:
    def __init__(self, path, mode="r"):
        """mode "r" (read), "r+" (read-write) or "w" (create, truncating)."""
        self.h = C.c_void_p()
        p = os.fsencode(path)
        rc = lib.adga_h5_create(p, C.byref(self.h)) if mode == "w" else lib.adga_h5_open(p, int(mode == "r+"),
                                                                                          C.byref(self.h))
        if rc != 0:
            raise H5Error(rc, f"cannot open {path}")

    def close(self):
        if self.h:
            lib.adga_h5_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            buf = C.create_string_buffer(512)
            lib.adga_h5_last_error(self.h, buf, 512)
            raise H5Error(rc, buf.value.decode(errors="replace"))

    # ------------------------------------------------------------------ 1P file
    def read_siw(self, dmft_iter, ndim, ndims, iwmax, orb_sym=False):
        out = np.zeros((2 * iwmax, ndim), dtype=np.complex128, order="F")
        at, keep = _atoms(ndim, ndims)
        self._check(lib.adga_h5_read_siw(self.h, dmft_iter.encode(), C.byref(at), int(orb_sym), iwmax, _p(out)))
        return out

    def read_giw(self, dmft_iter, ndim, ndims, iwmax, orb_sym=False):
        out = np.zeros((2 * iwmax, ndim), dtype=np.complex128, order="F")
        at, keep = _atoms(ndim, ndims)
        self._check(lib.adga_h5_read_giw(self.h, dmft_iter.encode(), C.byref(at), int(orb_sym), iwmax, _p(out)))
        return out

    def read_dc(self, dmft_iter, ndim, ndims):
        out = np.zeros((2, ndim), dtype=np.complex128, order="F")
        at, keep = _atoms(ndim, ndims)
        self._check(lib.adga_h5_read_dc(self.h, dmft_iter.encode(), C.byref(at), _p(out)))
        return out

    def read_mu(self, dmft_iter):
        v = C.c_double()
        self._check(lib.adga_h5_read_mu(self.h, dmft_iter.encode(), C.byref(v)))
        return v.value

    def read_beta(self):
        v = C.c_double()
        self._check(lib.adga_h5_read_beta(self.h, C.byref(v)))
        return v.value

    def read_hk(self):
        n, nk = C.c_int(), C.c_int()
        self._check(lib.adga_h5_read_hk(self.h, C.byref(n), C.byref(nk), None))
        out = np.zeros((n.value, n.value, nk.value), dtype=np.complex128, order="F")
        self._check(lib.adga_h5_read_hk(self.h, C.byref(n), C.byref(nk), _p(out)))
        return out

    # ------------------------------------------------------------------ 2P and V(q) files
    def read_vertex(self, iwb, ndim, ndims, iwfmax, iwbmax, iwfmax_small, beta, giw, vertex_type):
        g = _z(giw)
        at, keep = _atoms(ndim, ndims)
        v = Vertex(at, iwfmax, iwbmax, iwfmax_small, g.shape[0] // 2, vertex_type, float(beta), _p(g))
        N = ndim * ndim * 2 * iwfmax_small
        d = np.zeros((N, N), dtype=np.complex128, order="F")
        m = np.zeros_like(d)
        self._check(lib.adga_h5_read_vertex(self.h, int(iwb), C.byref(v), _p(d), _p(m)))
        return d, m

    def write_vertex(self, ineq, iwb, iwbmax, g4_dens, g4_magn):
        """g4(nb, nb, 2*iwfmax, nb, nb, 2*iwfmax) of both channels as stored (before the factor beta)."""
        d, m = _z(g4_dens), _z(g4_magn)
        nb, nfb = d.shape[0], d.shape[2]
        self._check(lib.adga_h5_write_vertex(self.h, ineq, int(iwb), iwbmax, nb, nfb // 2, _p(d), _p(m)))

    def read_vq(self, ndim, nqp, q_vol=True, q_data=None):
        out = np.zeros((ndim * ndim, ndim * ndim, nqp), dtype=np.complex128, order="F")
        q = None if q_data is None else np.ascontiguousarray(q_data, dtype=np.int32)
        self._check(lib.adga_h5_read_vq(self.h, ndim, nqp, int(q_vol), None if q is None else q.ctypes.data_as(_ip),
                                        _p(out)))
        return out

    # ------------------------------------------------------------------ output file
    def init_output(self, **kw):
        """Keywords: the fields of adga_h5_run (ints, mu, beta) and the arrays dc, giw, siw, hk, n_dmft,
        n_fock, k_data, q_data, k_data_eom."""
        arrays = {}
        r = Run()
        for name, ctype in Run._fields_:
            if name not in kw or kw[name] is None:
                continue
            if ctype is _dp:
                a = np.asfortranarray(kw[name], dtype=np.float64 if name == "k_data" else np.complex128)
                arrays[name] = a
                setattr(r, name, a.ctypes.data_as(_dp))
            elif ctype is _ip:
                a = np.ascontiguousarray(kw[name], dtype=np.int32)
                arrays[name] = a
                setattr(r, name, a.ctypes.data_as(_ip))
            else:
                setattr(r, name, kw[name])
        self._check(lib.adga_h5_init_output(self.h, C.byref(r)))

    def write_chi_qw(self, channel, chi_qw, ndim, iwbmax_small, nq, full=True, gzip=4):
        c = _z(chi_qw)
        nq = (C.c_int * 3)(*nq)
        self._check(lib.adga_h5_write_chi_qw(self.h, channel.encode(), int(full), ndim, iwbmax_small,
                                             C.cast(nq, _ip), _p(c), gzip))

    def write_chi_qpath(self, channel, chi, ndim, iwbmax_small, nqp, full=True):
        c = _z(chi)
        self._check(lib.adga_h5_write_chi_qpath(self.h, channel.encode(), int(full), ndim, iwbmax_small, nqp, _p(c)))

    def write_chi_loc(self, channel, chi_loc, ndim, iwbmax_small, full=True, gzip=4):
        c = _z(chi_loc)
        self._check(lib.adga_h5_write_chi_loc(self.h, channel.encode(), int(full), ndim, iwbmax_small, _p(c), gzip))

    def write_eom(self, sigma_sum_dmft, ndim, iwfmax_small, nk, sigma_sum=None, sigma_sum_hf=None, sigma_loc=None,
                  nonlocal_=True, gzip=4):
        arrs = [None if a is None else _z(a) for a in (sigma_sum, sigma_sum_hf, sigma_loc, sigma_sum_dmft)]
        nkc = (C.c_int * 3)(*nk)
        self._check(lib.adga_h5_write_eom(self.h, int(nonlocal_), ndim, iwfmax_small, C.cast(nkc, _ip),
                                          *[_p(a) for a in arrs], gzip))

    def write_occ(self, n_dga, n_dga_k, ndim, nk):
        a, b = _z(n_dga), _z(n_dga_k)
        nkc = (C.c_int * 3)(*nk)
        self._check(lib.adga_h5_write_occ(self.h, ndim, C.cast(nkc, _ip), _p(a), _p(b)))

    # ------------------------------------------------------------------ generic access
    def dims(self, path):
        r = C.c_int()
        d = (C.c_longlong * 8)()
        self._check(lib.adga_h5_dims(self.h, path.encode(), C.byref(r), d))
        return tuple(d[i] for i in range(r.value))

    def read_complex(self, path):
        """The whole compound (r, i) dataset in file (C) order."""
        out = np.zeros(self.dims(path), dtype=np.complex128)
        self._check(lib.adga_h5_read_complex(self.h, path.encode(), out.ctypes.data_as(_dp)))
        return out

    def read_real(self, path):
        out = np.zeros(self.dims(path), dtype=np.float64)
        self._check(lib.adga_h5_read_real(self.h, path.encode(), out.ctypes.data_as(_dp)))
        return out

    def write_complex(self, path, a):
        """A compound (r, i) dataset holding the C-order array a (its shape is the dataset's dims)."""
        a = np.ascontiguousarray(a, dtype=np.complex128)
        d = (C.c_longlong * max(1, a.ndim))(*a.shape)
        self._check(lib.adga_h5_write_complex(self.h, path.encode(), a.ndim, d, a.ctypes.data_as(_dp)))

    def write_real(self, path, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = (C.c_longlong * max(1, a.ndim))(*a.shape)
        self._check(lib.adga_h5_write_real(self.h, path.encode(), a.ndim, d, a.ctypes.data_as(_dp)))

    def write_attr(self, group, name, value):
        self._check(lib.adga_h5_write_attr_real(self.h, group.encode(), name.encode(), float(value)))
