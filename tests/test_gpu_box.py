"""A fermionic box that shrinks with |omega| (the 5-orbital workload's 2 x (80 - 2|omega|) frequencies): one
context with the largest box, every omega solved at its own BSE size through adga_set_local_omega_box.  The
reference has one global iwfmax_small, so each omega slice is pinned by the oracle run with that slice's box
as iwfmax_small: chi per point directly, and the self-energies as the sum of the slices' contributions, each
embedded in the largest box (a slice adds nothing at the frequencies outside its own box)."""
import numpy as np
import pytest

from conftest import rel_err
from adga_amd import synth

pytestmark = pytest.mark.gpu


def _slices(name, points, box_fn, **overrides):
    """Per-omega oracle runs with the slice's box, the self-energies embedded in the largest box."""
    import adga_oracle
    wl = synth.workload(name, **overrides)
    iwmax = wl.n_iwmax
    ref = {"chi_d": {}, "chi_m": {}, "bubble_nl": {}}
    n = wl.n
    snl = np.zeros((n, n, wl.nf, wl.nkp), dtype=complex)
    sdm = np.zeros((n, n, wl.nf), dtype=complex)
    for w in sorted({p[0] for p in points}):
        box = box_fn(w)
        wls = synth.workload(name, iwfmax_small=box, iwmax=iwmax, **overrides)
        inps = synth.build_inputs(wls)
        pts = [p for p in points if p[0] == w]
        print(f"oracle slice omega={w} box={box} points={len(pts)}", flush=True)
        out = adga_oracle.HotPath(inps).run(pts, lambda ww, c: synth.chi_loc(inps, ww, c))
        for j, p in enumerate(pts):
            for k in ref:
                ref[k][p] = out[k][:, :, j]
        o = wl.iwfmax_small - box
        snl[:, :, o:o + 2 * box, :] += out["sigma_nl"]
        sdm[:, :, o:o + 2 * box] += out["sigma_dmft"]
    return ref, snl, sdm, iwmax


def _run(name, points, box_fn, precision, tol, modes=(None, None), **overrides):
    from adga_amd import api
    ref, snl, sdm, iwmax = _slices(name, points, box_fn, **overrides)
    wl = synth.workload(name, iwmax=iwmax, **overrides)
    inp = synth.build_inputs(wl)
    inputs = {}

    def chi_loc(w, c):  # chi_loc of the slice's box, as the oracle slice had it
        if w not in inputs:
            inputs[w] = synth.build_inputs(synth.workload(name, iwfmax_small=box_fn(w), iwmax=iwmax, **overrides))
        return synth.chi_loc(inputs[w], w, c)

    print(f"{name} boxes {precision}", flush=True)
    with api.Context(inp, bubble=modes[0], eom=modes[1]) as ctx:
        out = api.run_loop(ctx, points, chi_loc, precision, box_fn=box_fn)
    assert (out["status"] == 0).all()
    for k in ("chi_d", "chi_m", "bubble_nl"):
        got = out[k]
        exp = np.stack([ref[k][p] for p in points], axis=2)
        assert rel_err(got, exp) < tol, (k, rel_err(got, exp))
    assert rel_err(out["sigma_nl"], snl) < tol, rel_err(out["sigma_nl"], snl)
    assert rel_err(out["sigma_dmft"], sdm) < tol, rel_err(out["sigma_dmft"], sdm)


@pytest.mark.parametrize("modes", [(1, 1), (0, 0)], ids=["dft", "direct"])
@pytest.mark.parametrize("precision", ["fp64", "refined"])
def test_gpu_shrinking_box_small(precision, modes):
    """3-orbital 4^3 workload, box 6 - |omega| over omega in [-2, 2] (N = 108 ... 72), every third q; the
    bubble and EOM as k-grid convolutions or direct sums."""
    box_fn = lambda w: 6 - abs(w)  # noqa: E731
    points = [(w, q) for w in (0, 1, -2, 2) for q in range(0, 64, 3)]
    _run("t2g_k4", points, box_fn, precision, 1e-10, modes)


@pytest.mark.parametrize("precision", ["fp64", "refined"])
def test_gpu_shrinking_box_five_orbitals(precision):
    """5-orbital d-shell model on 16^3 k, box 2 x (80 - 2|omega|): omega = 0, 5 and -20 at N = 4000, 3500
    and 2000 in one context, one point each (q = Gamma at omega = 0, so the local EOM runs too)."""
    box_fn = lambda w: 80 - 2 * abs(w)  # noqa: E731
    points = [(0, 0), (5, 1234), (-20, 77)]
    _run("d5_k16", points, box_fn, precision, 1e-10)


@pytest.mark.parametrize("precision", ["fp64", "refined"])
def test_gpu_shrinking_box_with_large_bubble_range(precision):
    """The box shrinking with |omega| together with the bubble over the large frequency range (the debug
    option "Bubble", bubble_range = 1): each omega's outer frequencies are the large range outside that
    omega's own box; every slice against the oracle run with the slice's box and the same range."""
    box_fn = lambda w: 6 - abs(w)  # noqa: E731
    points = [(w, q) for w in (0, 1, -2) for q in range(0, 64, 5)]
    _run("t2g_k4", points, box_fn, precision, 1e-10, bubble_range=1)
