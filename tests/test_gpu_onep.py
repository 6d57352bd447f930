"""Parity of the one-particle routines on the GPU (adga_get_giw, adga_get_nfock,
adga_get_sigma_g_loc, adga_get_ndga, through the C ABI) with the reference's own
one_particle_quant_module (golden vectors, tests/golden/make_golden_onep.py) and with the
numpy oracle, including a reordered subset of the k grid as the EOM k list and the
whole t2g_k20 grid at 800 frequencies.  Bound: 1e-12 relative (max-abs normalised)."""
import dataclasses

import numpy as np
import pytest

from conftest import load_golden, rel_err
from adga_amd import synth
import adga_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12
CASES = ["onep_sc1_k8", "onep_two_k4", "onep_t2g_k4", "onep_d5_k4", "onep_t2g_k4_sub"]


@pytest.mark.parametrize("name", CASES)
def test_gpu_one_particle_matches_reference(name):
    from adga_amd import api
    from golden.make_golden_onep import fixture_inputs
    gold = load_golden(name)
    inp, ss = fixture_inputs(name, gold)
    print(f"one-particle {name}", flush=True)
    with api.Context(inp) as ctx:
        assert rel_err(ctx.get_giw(), gold["giw"]) < TOL
        assert rel_err(ctx.get_nfock(), gold["n_fock"]) < TOL
        ndk, nd = ctx.get_ndga(ss)
        assert rel_err(ndk, gold["n_dga_k"]) < TOL
        assert rel_err(nd, gold["n_dga"]) < TOL
        if len(inp.k_data_eom) == inp.hk.shape[2]:
            sl, gl = ctx.get_sigma_g_loc(ss)
            assert rel_err(sl, gold["sigma_loc"]) < TOL
            assert rel_err(gl, gold["gloc"]) < TOL
        else:  # the reference calls it on the whole grid only
            with pytest.raises(api.AdgaError) as e:
                ctx.get_sigma_g_loc(np.zeros((inp.wl.n, inp.wl.n, inp.wl.nf, inp.hk.shape[2]), complex))
            assert e.value.code == 1  # ADGA_E_ARG


def test_gpu_one_particle_full_grid_against_oracle():
    """t2g_k20 (8000 k points) with iwmax = 800, a random self-energy in the vertex box and a
    random dc, against the oracle."""
    from adga_amd import api
    wl = synth.workload("t2g_k20", iwmax=800)
    inp = synth.build_inputs(wl)
    rng = np.random.default_rng(11)
    n, nkp = wl.n, 8000
    inp = dataclasses.replace(inp, dc=0.1 * (rng.normal(size=(2, n)) + 1j * rng.normal(size=(2, n))))
    ss = 0.2 * (rng.normal(size=(n, n, wl.nf, nkp)) + 1j * rng.normal(size=(n, n, wl.nf, nkp)))
    print("one-particle t2g_k20 iwmax 800", flush=True)
    with api.Context(inp) as ctx:
        assert rel_err(ctx.get_giw(), adga_oracle.get_giw(inp)) < TOL
        assert rel_err(ctx.get_nfock(), adga_oracle.get_nfock(inp)) < TOL
        ndk, nd = ctx.get_ndga(ss)
        ondk, ond = adga_oracle.get_ndga(inp, ss)
        assert rel_err(ndk, ondk) < TOL and rel_err(nd, ond) < TOL
        sl, gl = ctx.get_sigma_g_loc(ss)
        osl, ogl = adga_oracle.get_sigma_g_loc(inp, ss)
        assert rel_err(sl, osl) < TOL and rel_err(gl, ogl) < TOL
