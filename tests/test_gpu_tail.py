"""The bubble beyond the vertex box on the GPU: the large frequency range (option bubble_range = 1, the
reference's debug option "Bubble"; its golden vectors are in test_gpu_parity and test_gpu_dropin) and the
Matsubara tail of the local bubble (bubble_range = 2, no reference counterpart).

* bubble_range = 2 on the 3-orbital 4^3 workload against the oracle: bubble_loc with the tail, bubble_nl
  with the outer frequencies, chi and Sigma, both precisions, both bubble modes (the tail's frequency sums
  are summed directly on the device and formed by partial fractions in the oracle).
* The flat band of tests/test_oracle_tail.py on the device: the tail-corrected bubble_loc against the
  Lindhard function, with the same K^-5 convergence.
* Changing the range asks for the omega's local data again."""
import numpy as np
import pytest

from conftest import rel_err
from adga_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bubble", [1, 0], ids=["dft", "direct"])
@pytest.mark.parametrize("precision", ["fp64", "refined"])
def test_gpu_tail_matches_oracle(precision, bubble):
    import adga_oracle
    from adga_amd import api
    wl = synth.workload("t2g_k4", bubble_range=2, iwmax=24)
    inp = synth.build_inputs(wl)
    points = [(w, q) for w in range(-2, 3) for q in range(0, 64, 5)]
    print(f"oracle t2g_k4 bubble_range=2, {len(points)} points", flush=True)
    ref = adga_oracle.HotPath(inp).run(points, lambda w, c: synth.chi_loc(inp, w, c))
    print(f"GPU {precision} bubble={bubble}", flush=True)
    with api.Context(inp, bubble=bubble) as ctx:
        out = api.run_loop(ctx, points, lambda w, c: synth.chi_loc(inp, w, c), precision)
    assert (out["status"] == 0).all()
    for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
        assert rel_err(out[k], ref[k]) < 1e-10, (k, rel_err(out[k], ref[k]))
    for w in ref["bubble_loc"]:
        assert rel_err(out["bubble_loc"][w], ref["bubble_loc"][w]) < 1e-12, w


def test_gpu_tail_converges_to_the_lindhard_function():
    from adga_amd import api
    from test_oracle_tail import flat_band_inputs, lindhard
    errs = []
    for iwmax in (32, 64):
        inp, a = flat_band_inputs(iwmax)
        K = iwmax - inp.wl.iwbmax_small
        worst = 0.0
        print(f"flat band iwmax={iwmax}", flush=True)
        with api.Context(inp) as ctx:
            for w in range(-inp.wl.iwbmax_small, inp.wl.iwbmax_small + 1):
                _, _, bl = ctx.set_local_omega(w, synth.chi_loc(inp, w, "dens"), synth.chi_loc(inp, w, "magn"))
                worst = max(worst, np.abs(bl - lindhard(a, inp.beta, w)).max())
        errs.append((K, worst))
    (K1, e1), (K2, e2) = errs
    assert e1 / e2 > (K2 / K1) ** 4.5 and e2 < 1e-6, errs


def test_gpu_bubble_range_change_needs_the_local_data_again():
    from adga_amd import api
    wl = synth.workload("t2g_k4", iwmax=20)
    inp = synth.build_inputs(wl)
    xd, xm = synth.chi_loc(inp, 1, "dens"), synth.chi_loc(inp, 1, "magn")
    with api.Context(inp) as ctx:
        _, _, box = ctx.set_local_omega(1, xd, xm)
        ctx.run_points(np.array([1, 2]), "fp64")
        lib = api.lib
        assert lib.adga_set_option(ctx.h, b"bubble_range", 1.0) == 0
        r = ctx.run_points(np.array([1, 2]), "fp64", raise_on_error=False)
        assert r["rc"] == api._lib.ADGA_E_STATE
        _, _, big = ctx.set_local_omega(1, xd, xm)
        assert np.abs(big - box).max() > 1e-6  # the large range adds frequencies
        assert lib.adga_set_option(ctx.h, b"bubble_range", 3.0) != 0


@pytest.mark.timeout(600)
def test_gpu_tail_full_size_sampled():
    """The benchmark workload (3-orbital t2g, 20^3 k, 2x60 x 31, N = 1080, iwmax = 300: 450 outer
    frequencies) with bubble_range = 2, two points at two omegas against the oracle, both precisions."""
    import adga_oracle
    from adga_amd import api
    wl = synth.workload("t2g_k20", bubble_range=2)
    inp = synth.build_inputs(wl)
    pts = [(0, 0), (3, 4321)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    print("oracle t2g_k20 bubble_range=2, 2 points", flush=True)
    ref = adga_oracle.HotPath(inp).run(pts, cl)
    for precision in ("fp64", "refined"):
        print(f"GPU {precision}", flush=True)
        with api.Context(inp) as ctx:
            out = api.run_loop(ctx, pts, cl, precision)
        assert (out["status"] == 0).all()
        for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
            assert rel_err(out[k], ref[k]) < 1e-10, (precision, k, rel_err(out[k], ref[k]))
        for w in ref["bubble_loc"]:
            assert rel_err(out["bubble_loc"][w], ref["bubble_loc"][w]) < 1e-12, w
