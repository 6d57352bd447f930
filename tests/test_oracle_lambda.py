"""Pins of the oracle's lambda correction (no reference counterpart; the physical susceptibility takes the
entries of the reference's reduced output, hdf5_module.f90:1479-1495).

* chi_phys picks chi[(i,i), (j,j)]: a susceptibility whose entries encode their indices.
* Closed forms of the sum rule: one omega with the same chi = c at every q gives S(lambda) = c / (beta
  (1 + lambda c)), so lambda = 1/(beta t) - 1/c for target t; a lattice chi equal to the local one at every q
  gives lambda = 0 for the local target.
* dS/dlambda against a four-point central difference; S decreasing from the pole on."""
import numpy as np
import pytest

import adga_oracle as O


def test_chi_phys_takes_the_reduced_entries():
    n = 3
    chi = np.zeros((n * n, n * n, 2), dtype=complex)
    for a in range(n * n):
        for b in range(n * n):
            chi[a, b, :] = 1000 * a + b + 1j * (a == b)
    exp = sum(chi[i * n + i, j * n + j, 0] for i in range(n) for j in range(n))
    assert np.allclose(O.chi_phys(chi), exp)
    # i1 = i*n + j with i != j never enters
    chi2 = chi.copy()
    chi2[1, 1, :] += 77.0
    assert np.allclose(O.chi_phys(chi2), O.chi_phys(chi))


@pytest.mark.parametrize("t_rel", [0.3, 0.9, 1.0, 2.5])
def test_lambda_closed_form_constant_chi(t_rel):
    beta, nqp, c = 7.0, 50, 0.8
    chi = np.full(nqp, c, dtype=complex)
    t = t_rel * c / beta
    lam = O.lambda_solve(chi, beta, nqp, t)
    assert abs(lam - (1 / (beta * t) - 1 / c)) <= 1e-12 * max(1.0, abs(lam))


def test_lambda_zero_when_lattice_equals_local():
    rng = np.random.default_rng(1)
    beta, nqp, nw = 10.0, 27, 5
    loc = 0.05 + 0.2 * rng.random(nw)
    loc[nw // 2] = 0.4
    chi = np.repeat(loc, nqp).astype(complex)
    t = loc.sum() / beta
    lam = O.lambda_solve(chi, beta, nqp, t, chi0=chi[(nw // 2) * nqp:(nw // 2 + 1) * nqp])
    assert abs(lam) < 1e-13


def test_lambda_derivative_and_monotonic():
    rng = np.random.default_rng(2)
    beta, nqp = 10.0, 64
    chi = 0.05 + 0.3 * rng.random(3 * nqp) + 1e-3j * rng.normal(size=3 * nqp)
    h = 1e-3
    for lam in (-1.5, 0.0, 0.7, 4.0):
        s = lambda x: O.lambda_sums(chi, beta, nqp, x)[0]  # noqa: E731
        fd = (8 * (s(lam + h) - s(lam - h)) - (s(lam + 2 * h) - s(lam - 2 * h))) / (12 * h)
        ds = O.lambda_sums(chi, beta, nqp, lam)[1]
        assert abs(ds - fd) <= 1e-9 * abs(ds)
    vals = [O.lambda_sums(chi, beta, nqp, x)[0].real for x in np.linspace(-1 / chi.real.max() + 0.05, 10, 40)]
    assert np.all(np.diff(vals) < 0)
