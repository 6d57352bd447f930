"""Writes the golden vectors of the one-particle routines (get_giw, get_nfock, get_sigma_g_loc,
get_ndga; src/one_particle_quant_module.f90) under tests/golden/onep_*.npz from the reference's
own Fortran module (oracle/_ref/adga_op_harness, built by `make -C oracle` against the reference
sources with amdflang and LAPACK/BLAS from MKL 2021.4).

The one-particle inputs hk and siw come from the seeded generator adga_amd.synth (a digest is
stored); dc, the self-energy sigma_sum and, for one case, a reordered subset of the k grid as the
EOM k list are drawn here and stored with the outputs.

Usage: python tests/golden/make_golden_onep.py [name ...]
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from adga_amd import synth  # noqa: E402
import ref_runner  # noqa: E402

# name -> (workload, EOM k list: "all" = the whole grid, "sub" = every third k in reverse order)
FIXTURES = {
    "onep_sc1_k8": ("sc1_k8", "all"),
    "onep_two_k4": ("two_k4", "all"),
    "onep_t2g_k4": ("t2g_k4", "all"),
    "onep_d5_k4": ("d5_k4", "all"),
    "onep_t2g_k4_sub": ("t2g_k4", "sub"),
}


def inputs_digest(inp):
    h = hashlib.sha256()
    for a in (inp.hk, inp.siw):
        h.update(np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).tobytes())
    return h.hexdigest()


def fixture_inputs(name, stored=None):
    """The workload's inputs with this fixture's dc and k list, and its sigma_sum: drawn from a
    seed when `stored` is None, else taken from the stored fixture."""
    wname, kl = FIXTURES[name]
    wl = synth.workload(wname)
    inp = synth.build_inputs(wl)
    n, nkp = wl.n, int(np.prod(wl.nk))
    if stored is None:
        rng = np.random.default_rng(sum(map(ord, name)))
        dc = 0.1 * (rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n)))
        keom = np.arange(1, nkp + 1, dtype=np.int32) if kl == "all" else np.arange(nkp, 0, -3, dtype=np.int32)
        shape = (n, n, wl.nf, len(keom))
        ss = 0.2 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    else:
        dc, keom, ss = stored["dc"], stored["k_data_eom"], stored["sigma_sum"]
    inp = dataclasses.replace(inp, dc=np.asarray(dc), k_data_eom=np.asarray(keom, dtype=np.int32))
    return inp, np.asarray(ss)


def make(name):
    inp, ss = fixture_inputs(name)
    ref = ref_runner.run_one_particle(inp, ss)
    out = {k: v for k, v in ref.items() if k != "wall_s"}
    out.update(dc=inp.dc, k_data_eom=inp.k_data_eom, sigma_sum=ss, digest=np.array(inputs_digest(inp)),
               provenance=np.array("reference module src/one_particle_quant_module.f90 (get_giw, get_nfock, "
                                   "get_sigma_g_loc, get_ndga) built by amdflang 22 (ROCm 7.2), "
                                   "LAPACK = MKL 2021.4"))
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: {os.path.getsize(path) / 1e3:.0f} kB, ref {ref['wall_s']:.2f}s")


if __name__ == "__main__":
    if not ref_runner.op_available():
        sys.exit("oracle/_ref/adga_op_harness is not built (make -C oracle)")
    for nm in (sys.argv[1:] or FIXTURES):
        make(nm)
