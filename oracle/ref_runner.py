"""Drives oracle/_ref/adga_ref_harness (the reference's own Fortran hot-path
modules plus a test-only loop driver) on a synthetic workload.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to write golden
vectors, by tests that compare the numpy oracle with the reference where the
binary exists, and by bench.py's `cpu_baseline` leg.
"""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "adga_ref_harness")


def available() -> bool:
    return os.path.exists(HARNESS)


def _w(f, a, dtype):
    np.asarray(a, dtype=dtype).flatten(order="F").tofile(f)


def run(inp, points, chi_loc_fn, threads=1, timeout=3600, exe=None, args=()):
    """Run the reference modules over `points` = [(iwb, iq0), ...] (iq0 0-based).
    Returns a dict of Fortran-shaped complex128 arrays and the wall time.  `exe` runs another
    program that reads and writes the same file layout (the GPU path's Fortran host check)."""
    wl = inp.wl
    n, n2, nf = wl.n, wl.n * wl.n, wl.nf
    N = n2 * nf
    nw = 2 * wl.iwbmax_small + 1
    npts = len(points)
    nkp_eom = len(inp.k_data_eom)
    wused = sorted({p[0] for p in points})
    # header flags: bit 0 V(q), bit 1 EOM k list other than the grid in order, bit 2 no EOM (q path),
    # bit 3 the bubble over the large range (debug option "Bubble"); the tail has no reference counterpart
    kpath = nkp_eom != int(np.prod(wl.nk)) or not np.array_equal(inp.k_data_eom, np.arange(1, nkp_eom + 1))
    br = getattr(wl, "bubble_range", 0)
    if br == 2:
        raise ValueError("the reference has no Matsubara tail (bubble_range 2)")
    flags = (1 if inp.v_q is not None else 0) | (2 if kpath else 0) | (0 if wl.do_eom else 4) | (8 if br else 0)
    with tempfile.TemporaryDirectory() as td:
        fi = os.path.join(td, "in.bin")
        fo = os.path.join(td, "out.bin")
        with open(fi, "wb") as f:
            _w(f, [n, *wl.nk, inp.iwmax, wl.iwfmax_small, wl.iwbmax_small, len(inp.q_data),
                   nkp_eom, npts, flags], np.int32)
            _w(f, [inp.beta, inp.mu], np.float64)
            for a in (inp.hk, inp.siw, inp.giw, inp.dc, inp.u, inp.u_tilde, inp.n_dmft):
                _w(f, a, np.complex128)
            _w(f, inp.q_data, np.int32)
            _w(f, inp.k_data_eom, np.int32)
            if inp.v_q is not None:
                _w(f, inp.v_q, np.complex128)
            _w(f, [p[0] for p in points], np.int32)
            _w(f, [p[1] + 1 for p in points], np.int32)
            _w(f, [len(wused)], np.int32)
            _w(f, wused, np.int32)
            for iwb in wused:
                _w(f, chi_loc_fn(iwb, "dens"), np.complex128)
                _w(f, chi_loc_fn(iwb, "magn"), np.complex128)
        env = dict(os.environ)
        env.update(MKL_NUM_THREADS=str(threads), MKL_DYNAMIC="FALSE", OMP_NUM_THREADS=str(threads))
        import time
        t0 = time.perf_counter()
        subprocess.run([exe or HARNESS, fi, fo, *args], check=True, env=env, timeout=timeout)
        dt = time.perf_counter() - t0
        raw = np.fromfile(fo, dtype=np.complex128)
    shapes = [("chi_d", (n2, n2, npts)), ("chi_m", (n2, n2, npts)), ("bubble_nl", (n2, n2, npts)),
              ("sigma_nl", (n, n, nf, nkp_eom)), ("sigma_dmft", (n, n, nf)),
              ("chi_loc_d", (n2, n2, nw)), ("chi_loc_m", (n2, n2, nw)), ("bubble_loc", (n2, n2, nw)),
              ("gammaqd0", (n2, N)), ("eta_d0", (n2, N)), ("eta_m0", (n2, N)), ("sigma_hf", (n, n, nkp_eom))]
    out = {}
    o = 0
    for name, shp in shapes:
        sz = int(np.prod(shp))
        out[name] = raw[o:o + sz].reshape(shp, order="F")
        o += sz
    assert o == raw.size, (o, raw.size)
    out["wall_s"] = dt
    return out


OP_HARNESS = os.path.join(HERE, "_ref", "adga_op_harness")


def op_available() -> bool:
    return os.path.exists(OP_HARNESS)


def run_one_particle(inp, sigma_sum, timeout=3600):
    """Run the reference's get_giw, get_nfock, get_sigma_g_loc (when nkp_eom == nkp) and get_ndga
    (oracle/_ref/adga_op_harness) on `inp` and sigma_sum (n, n, 2*iwfmax_small, nkp_eom).
    Returns a dict of Fortran-shaped complex128 arrays and the wall time."""
    wl = inp.wl
    n, nf = wl.n, wl.nf
    nkp = int(np.prod(wl.nk))
    nkp_eom = len(inp.k_data_eom)
    with tempfile.TemporaryDirectory() as td:
        fi = os.path.join(td, "in.bin")
        fo = os.path.join(td, "out.bin")
        with open(fi, "wb") as f:
            _w(f, [n, *wl.nk, inp.iwmax, wl.iwfmax_small, nkp_eom], np.int32)
            _w(f, [inp.beta, inp.mu], np.float64)
            for a in (inp.hk, inp.siw, inp.dc):
                _w(f, a, np.complex128)
            _w(f, inp.k_data_eom, np.int32)
            _w(f, sigma_sum, np.complex128)
        env = dict(os.environ)
        env.update(MKL_NUM_THREADS="1", MKL_DYNAMIC="FALSE", OMP_NUM_THREADS="1")
        import time
        t0 = time.perf_counter()
        subprocess.run([OP_HARNESS, fi, fo], check=True, env=env, timeout=timeout)
        dt = time.perf_counter() - t0
        raw = np.fromfile(fo, dtype=np.complex128)
    shapes = [("giw", (2 * inp.iwmax, n)), ("n_fock", (nkp, n, n)), ("sigma_loc", (n, n, nf)),
              ("gloc", (2 * inp.iwmax, n, n)), ("n_dga_k", (nkp_eom, n, n)), ("n_dga", (n,))]
    out = {}
    o = 0
    for name, shp in shapes:
        sz = int(np.prod(shp))
        out[name] = raw[o:o + sz].reshape(shp, order="F")
        o += sz
    assert o == raw.size, (o, raw.size)
    out["wall_s"] = dt
    return out
