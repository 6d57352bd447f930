"""Pins the numpy oracle to golden vectors written by the reference's own Fortran
hot-path modules (tests/golden/make_golden.py, oracle/_ref).  Bound: 1e-10 relative
(max-abs normalised) for every accumulated quantity."""
import numpy as np
import pytest

from conftest import load_golden, rel_err
from adga_amd import synth
import adga_oracle
from golden.make_golden import FIXTURES, input_digest

CASES = ["sc1_k4", "two_k4", "t2g_k4", "d5_k4", "sc1_k8"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference(name):
    gold = load_golden(name)
    wl = synth.workload(FIXTURES[name][0])
    inp = synth.build_inputs(wl)
    pts = [tuple(p) for p in gold["points"].tolist()]
    assert str(gold["digest"]) == input_digest(inp, pts), "synthetic input generator changed"
    hp = adga_oracle.HotPath(inp)
    out = hp.run(pts, lambda w, c: synth.chi_loc(inp, w, c), keep={pts[0]})
    for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
        assert rel_err(out[k], gold[k]) < 1e-10, k
    p0 = out["points"][pts[0]]
    assert rel_err(p0["gammaqd"], gold["gammaqd0"]) < 1e-10
    assert rel_err(p0["eta_d"], gold["eta_d0"]) < 1e-10
    assert rel_err(p0["eta_m"], gold["eta_m0"]) < 1e-10
    for w in out["chi_loc_d"]:
        i = w + wl.iwbmax_small
        assert rel_err(out["chi_loc_d"][w], gold["chi_loc_d"][:, :, i]) < 1e-10
        assert rel_err(out["chi_loc_m"][w], gold["chi_loc_m"][:, :, i]) < 1e-10
        assert rel_err(out["bubble_loc"][w], gold["bubble_loc"][:, :, i]) < 1e-12


def test_golden_fixture_files_are_small():
    import os
    from conftest import GOLDEN
    tot = sum(os.path.getsize(os.path.join(GOLDEN, f)) for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert tot < 8e6
