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


def test_gpu_get_giw_replaces_the_context_giw():
    """A context created with giw = 0 gets the lattice giw from adga_get_giw, as get_giw
    overwrites the module's giw (main.F90:207); the loop then matches the reference's golden
    vectors (whose giw is the same lattice sum), and after another adga_get_giw the next
    adga_run_points needs adga_set_local_omega first (ADGA_E_STATE)."""
    from adga_amd import api
    from golden.make_golden import FIXTURES
    gold = load_golden("t2g_k4")
    wl = synth.workload(FIXTURES["t2g_k4"][0])
    inp = synth.build_inputs(wl)
    pts = [tuple(p) for p in gold["points"].tolist()]
    print("one-particle get_giw into the context (t2g_k4 loop)", flush=True)
    with api.Context(dataclasses.replace(inp, giw=np.zeros_like(inp.giw))) as ctx:
        assert rel_err(ctx.get_giw(), inp.giw) < 1e-13
        out = api.run_loop(ctx, pts, lambda w, c: synth.chi_loc(inp, w, c), "fp64")
        for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
            assert rel_err(out[k], gold[k]) < 1e-10, k
        ctx.get_giw()
        with pytest.raises(api.AdgaError) as e:
            ctx.run_points(np.array([1], dtype=np.int32))
        assert e.value.code == 6  # ADGA_E_STATE


def test_gpu_sigma_sum_of_the_context():
    """adga_get_sigma_g_loc / adga_get_ndga with sigma_sum = NULL take the context's own: the Sigma_nl it
    accumulated over the t2g_k4 loop plus siw on the orbital diagonal, formed on the device as
    add_siw_dmft forms it (src/eom_module.f90:237-251).  Equal, exactly, to the same routines given that
    array from the host, also on a second call."""
    from adga_amd import api
    from golden.make_golden import FIXTURES
    gold = load_golden("t2g_k4")
    wl = synth.workload(FIXTURES["t2g_k4"][0])
    inp = synth.build_inputs(wl)
    pts = [tuple(p) for p in gold["points"].tolist()]
    n, iwf, iwmax = wl.n, wl.iwfmax_small, inp.siw.shape[0] // 2
    print("sigma_sum of the context (t2g_k4 loop)", flush=True)
    with api.Context(inp) as ctx:
        out = api.run_loop(ctx, pts, lambda w, c: synth.chi_loc(inp, w, c), "fp64")
        ss = np.array(out["sigma_nl"], order="F")
        for i in range(n):
            ss[i, i, :, :] += inp.siw[iwmax - iwf:iwmax + iwf, i][:, None]
        assert np.abs(out["sigma_nl"]).max() > 0
        host = ctx.get_sigma_g_loc(ss) + ctx.get_ndga(ss)
        for _ in range(2):
            own = ctx.get_sigma_g_loc() + ctx.get_ndga()
            for a, b in zip(host, own):
                assert np.array_equal(a, b)
