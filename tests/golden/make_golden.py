"""Writes the golden vectors under tests/golden/ from the reference's own Fortran
hot-path modules (oracle/_ref/adga_ref_harness, built by `make -C oracle`
against the reference sources with amdflang and LAPACK/BLAS from MKL 2021.4,
/opt/conda/lib/libmkl_rt.so).

Each fixture holds the reference outputs for a synthetic workload (inputs are
regenerated from the seeded generator adga_amd.synth; a digest of those inputs
is stored so that a changed generator is detected instead of silently compared).

Usage: python tests/golden/make_golden.py [name ...]
"""
from __future__ import annotations

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

# name -> (workload, max points or None for the whole (q, omega) list)
FIXTURES = {
    "sc1_k8": ("sc1_k8", None),     # 1-orbital simple cubic, 8^3, 2x20 x 11: all 5632 points
    "sc1_k4": ("sc1_k4", None),
    "two_k4": ("two_k4", None),
    "t2g_k4": ("t2g_k4", None),
    "d5_k4": ("d5_k4", 24),
    "t2g_k4_keom": ("t2g_k4_keom", None),    # EOM on a reordered third of the k grid, q on a 2^3 sub-grid
    "t2g_k6_qpath": ("t2g_k6_qpath", None),  # chi along a q path (no EOM)
    "two_k4_vq": ("two_k4_vq", None),        # V(q): density BSE term, Hartree/Fock, V terms of the EOM
    "t2g_k4_vq": ("t2g_k4_vq", None),        # V(q) with the EOM on a reordered k list
    "t2g_k4_bub": ("t2g_k4_bub", None),      # bubble over the large frequency range (debug option Bubble)
    "two_k4_bub": ("two_k4_bub", None),      # the same, with the EOM on a reordered k list
}


def input_digest(inp, points):
    h = hashlib.sha256()
    for a in (inp.hk, inp.siw, inp.giw, inp.u, inp.u_tilde, inp.n_dmft):
        h.update(np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).tobytes())
    # the q and EOM k lists and V(q) enter only when they differ from the defaults, so the digests of the
    # fixtures written before these inputs existed stay valid
    if inp.wl.bubble_range:
        h.update(np.int32(inp.wl.bubble_range).tobytes())
    if inp.wl.q_list != "vol" or inp.wl.k_eom != "all" or inp.v_q is not None:
        h.update(np.ascontiguousarray(inp.q_data, dtype=np.int32).tobytes())
        h.update(np.ascontiguousarray(inp.k_data_eom, dtype=np.int32).tobytes())
        if inp.v_q is not None:
            h.update(np.ascontiguousarray(inp.v_q).tobytes())
    for iwb in sorted({p[0] for p in points}):
        for ch in ("dens", "magn"):
            h.update(np.ascontiguousarray(synth.chi_loc(inp, iwb, ch)).tobytes())
    return h.hexdigest()


def points_for(wl, maxp):
    pts = [(w, q) for w in range(-wl.iwbmax_small, wl.iwbmax_small + 1) for q in range(wl.nqp)]
    return pts if maxp is None else pts[:maxp]


def make(name):
    wname, maxp = FIXTURES[name]
    wl = synth.workload(wname)
    inp = synth.build_inputs(wl)
    pts = points_for(wl, maxp)
    ref = ref_runner.run(inp, pts, lambda w, c: synth.chi_loc(inp, w, c))
    out = {k: v for k, v in ref.items() if k != "wall_s"}
    out["points"] = np.array(pts, dtype=np.int32)
    out["digest"] = np.array(input_digest(inp, pts))
    out["provenance"] = np.array("reference modules src/{one_particle_quant,susc,eom,lapack}_module.f90 "
                                 "built by amdflang 22 (ROCm 7.2), LAPACK/BLAS = MKL 2021.4")
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: {len(pts)} points, {os.path.getsize(path) / 1e3:.0f} kB, ref {ref['wall_s']:.1f}s")


if __name__ == "__main__":
    if not ref_runner.available():
        sys.exit("oracle/_ref/adga_ref_harness is not built (make -C oracle)")
    for nm in (sys.argv[1:] or FIXTURES):
        make(nm)
