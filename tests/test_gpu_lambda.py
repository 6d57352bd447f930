"""The lambda correction on the GPU (adga_lambda_*) against the oracle (tests/test_oracle_lambda.py pins it).

* Closed form: one omega with the same chi = c at every q, lambda = 1/(beta t) - 1/c for target t.
* The final magnetic and density susceptibilities of the 3-orbital 4^3 workload as the GPU path produces
  them (every (omega, q) point, chi + bubble_nl + chi_loc + bubble_loc): the device's S(lambda), dS/dlambda
  and the local target against the oracle's numpy sums (1e-13), lambda against the oracle's Brent root
  for the local target and for targets on both sides of S(0), the corrected chi_phys; repeated sums are
  bit-identical (fixed-order reductions).
* No positive omega = 0 susceptibility: refused at creation (host check)."""
import numpy as np
import pytest

from conftest import rel_err
from adga_amd import synth
import adga_oracle as O

pytestmark = pytest.mark.gpu


def _as_chi_qw(chi_phys_values, n=1):
    """(n^2, n^2, npts) with chi_phys = the given values (n = 1: the value itself)."""
    return np.asfortranarray(np.asarray(chi_phys_values, dtype=np.complex128).reshape(1, 1, -1))


@pytest.mark.parametrize("t_rel", [0.3, 1.0, 2.5])
def test_gpu_lambda_closed_form(t_rel):
    from adga_amd import api
    beta, nqp, c = 7.0, 50, 0.8
    chi = _as_chi_qw(np.full(nqp, c))
    loc = _as_chi_qw([c])
    with api.Lambda(chi, loc, beta, nqp) as lam:
        t = t_rel * c / beta
        x, it = lam.solve(target=t)
        assert abs(x - (1 / (beta * t) - 1 / c)) <= 1e-13 * max(1.0, abs(x)), (x, it)
        assert abs(lam.target() - c / beta) <= 1e-15


def _final_chi(out, wl, r):
    W, nq = wl.iwbmax_small, wl.nqp
    tot = out["chi_" + r] + out["bubble_nl"]
    for w in range(-W, W + 1):
        s = slice((w + W) * nq, (w + W + 1) * nq)
        tot[:, :, s] += (out["chi_loc_" + r][w] + out["bubble_loc"][w])[:, :, None]
    loc = np.stack([out["chi_loc_" + r][w] + out["bubble_loc"][w] for w in range(-W, W + 1)], axis=2)
    return np.asfortranarray(tot), np.asfortranarray(loc)


def test_gpu_lambda_on_the_final_susceptibilities():
    from adga_amd import api
    wl = synth.workload("t2g_k4")
    inp = synth.build_inputs(wl)
    W, nq = wl.iwbmax_small, wl.nqp
    points = [(w, q) for w in range(-W, W + 1) for q in range(nq)]
    print(f"t2g_k4: {len(points)} points on the GPU", flush=True)
    with api.Context(inp) as ctx:
        out = api.run_loop(ctx, points, lambda w, c: synth.chi_loc(inp, w, c), "fp64")
    for r in ("m", "d"):
        chi_qw, chi_loc = _final_chi(out, wl, r)
        cp = O.chi_phys(chi_qw)
        c0 = cp[W * nq:(W + 1) * nq]
        print(f"channel {r}: lambda sums and solves", flush=True)
        with api.Lambda(chi_qw, chi_loc, inp.beta, nq) as lam:
            t_loc = O.chi_phys(chi_loc).sum().real / inp.beta
            assert abs(lam.target() - t_loc) <= 1e-14 * abs(t_loc)
            pole = -1 / c0.real.max()
            xs = np.array([0.5 * pole, -1e-3, 0.0, 0.3, 2.0])
            s, ds = lam.sums(xs)
            for i, x in enumerate(xs):
                es, eds = O.lambda_sums(cp, inp.beta, nq, x)
                assert abs(s[i] - es) <= 1e-13 * abs(es) and abs(ds[i] - eds) <= 1e-13 * abs(eds), (x, s[i], es)
            s2, ds2 = lam.sums(xs)
            assert np.array_equal(s, s2) and np.array_equal(ds, ds2)
            s0 = O.lambda_sums(cp, inp.beta, nq, 0.0)[0].real
            for t in (t_loc, 0.5 * s0, 0.95 * s0, 1.5 * s0, 4.0 * s0):
                x, it = lam.solve(target=t)
                ex = O.lambda_solve(cp, inp.beta, nq, t, chi0=c0)
                assert abs(x - ex) <= 1e-11 * max(1.0, abs(ex)), (r, t, x, ex, it)
            x, _ = lam.solve()
            assert rel_err(lam.chi(x), cp / (1 + x * cp)) < 1e-14


def test_gpu_lambda_without_positive_static_susceptibility_is_refused():
    from adga_amd import api
    chi = _as_chi_qw(-np.ones(3 * 4))
    loc = _as_chi_qw(-np.ones(3))
    with pytest.raises(api.AdgaError) as e:
        api.Lambda(chi, loc, 5.0, 4)
    assert e.value.code == 1  # ADGA_E_ARG


def test_gpu_lambda_non_finite_input_is_reported():
    """A NaN target is refused; a non-finite susceptibility makes the solve report ADGA_E_NONFINITE
    instead of returning a lambda."""
    from adga_amd import api
    beta, nqp = 7.0, 10
    chi = np.full(3 * nqp, 0.8)
    chi[4] = np.nan  # an omega != 0 point: the pole is still bounded by the omega = 0 points
    with api.Lambda(_as_chi_qw(chi), _as_chi_qw(np.full(3, 0.8)), beta, nqp) as lam:
        with pytest.raises(api.AdgaError) as e:
            lam.solve(target=float("nan"))
        assert e.value.code == 1  # ADGA_E_ARG
        with pytest.raises(api.AdgaError) as e:
            lam.solve(target=0.1)
        assert e.value.code == 5  # ADGA_E_NONFINITE
