"""Parity of the HIP path (through the C ABI) with the reference.

* golden: the reference's own Fortran modules' outputs (tests/golden) on small workloads, every
  accumulated quantity, both precision modes.  Bound 1e-10 relative (max-abs normalised) in FP64
  and in refined mode (complex64 LU + FP64 refinement to a relative residual <= 1e-12, the
  default tolerance; one correction usually reaches it).
* full size: the 3-orbital t2g workload at 20^3 k, 2x60 x 31 frequencies (BSE N = 1080) against
  the numpy oracle on sampled points, and the F = 0 identity at that size.
"""
import numpy as np
import pytest

from conftest import load_golden, rel_err
from adga_amd import synth

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_REFINED = 1e-10


def _api():
    from adga_amd import api
    return api


def _compare_golden(out, gold, wl, tol):
    for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
        if not wl.do_eom and k.startswith("sigma"):
            assert np.abs(out[k]).max() == 0.0, k  # no EOM on a q path, as in the reference
            continue
        e = rel_err(out[k], gold[k])
        assert e < tol, (k, e)
    if "sigma_hf" in gold and np.abs(gold["sigma_hf"]).max() > 0:
        e = rel_err(out["sigma_hf"], gold["sigma_hf"])
        assert e < tol, ("sigma_hf", e)
    for w in out["chi_loc_d"]:
        i = w + wl.iwbmax_small
        assert rel_err(out["chi_loc_d"][w], gold["chi_loc_d"][:, :, i]) < tol
        assert rel_err(out["chi_loc_m"][w], gold["chi_loc_m"][:, :, i]) < tol
        assert rel_err(out["bubble_loc"][w], gold["bubble_loc"][:, :, i]) < tol


MODES = [(1, 1), (0, 0), (1, 0), (0, 1)]  # (bubble, eom): 1 = k-grid DFT, 0 = direct sums


@pytest.mark.parametrize("modes", MODES, ids=lambda m: f"bubble{m[0]}-eom{m[1]}")
@pytest.mark.parametrize("precision", ["fp64", "refined"])
@pytest.mark.parametrize("name", ["sc1_k4", "two_k4", "t2g_k4", "d5_k4", "sc1_k8", "t2g_k4_keom", "t2g_k6_qpath",
                                  "two_k4_vq", "t2g_k4_vq", "t2g_k4_bub", "two_k4_bub"])
def test_gpu_matches_reference_golden(name, precision, modes):
    from golden.make_golden import FIXTURES
    api = _api()
    gold = load_golden(name)
    wl = synth.workload(FIXTURES[name][0])
    inp = synth.build_inputs(wl)
    pts = [tuple(p) for p in gold["points"].tolist()]
    print(f"case {name} {precision} modes={modes} points={len(pts)}", flush=True)
    with api.Context(inp, batch=512, bubble=modes[0], eom=modes[1]) as ctx:
        out = api.run_loop(ctx, pts, lambda w, c: synth.chi_loc(inp, w, c), precision)
        st = ctx.stats()
    assert (out["status"] == 0).all()
    _compare_golden(out, gold, wl, TOL_FP64 if precision == "fp64" else TOL_REFINED)
    if precision == "refined":
        assert st["refine_resid"] <= 1e-12


def test_gpu_inverse_matrix_matches_numpy():
    api = _api()
    rng = np.random.default_rng(3)
    # sizes that reach every panel shape: rows <= 256 .. 1536 per 256-thread panel, > 1536 on 512
    for n in (1, 3, 40, 300, 700, 1300, 1700):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2 * np.eye(n)
        inv, ierr = api.inverse_matrix(a)
        assert ierr == 0
        assert rel_err(inv, np.linalg.inv(a)) < 1e-11


def test_gpu_inverse_matrix_singular_raises():
    api = _api()
    a = np.ones((8, 8), dtype=complex)
    with pytest.raises(api.AdgaError) as e:
        api.inverse_matrix(a)
    assert e.value.code == 3  # ADGA_E_SINGULAR, the reference's "ERROR in ZGETRF"


def test_gpu_nonfinite_vertex_is_reported():
    api = _api()
    wl = synth.workload("t2g_k4")
    inp = synth.build_inputs(wl)
    with api.Context(inp) as ctx:
        xd = synth.chi_loc(inp, 0, "dens")
        xd[3, 5] = np.nan
        ctx.set_local_omega(0, xd, synth.chi_loc(inp, 0, "magn"))
        r = ctx.run_points(np.array([1, 2, 3]), "fp64", raise_on_error=False)
        assert r["rc"] != 0 and (r["status"] != 0).all()
        r = ctx.run_points(np.array([1, 2]), "refined", raise_on_error=False)
        assert r["rc"] != 0 and (r["status"] != 0).all()


def test_gpu_state_and_argument_errors():
    api = _api()
    inp = synth.build_inputs(synth.workload("sc1_k4"))
    with api.Context(inp) as ctx:
        with pytest.raises(api.AdgaError) as e:
            ctx.run_points(np.array([1]))
        assert e.value.code == 6  # set_local_omega first
        ctx.set_local_omega(0, synth.chi_loc(inp, 0, "dens"), synth.chi_loc(inp, 0, "magn"))
        with pytest.raises(api.AdgaError) as e:
            ctx.run_points(np.array([0]))
        assert e.value.code == 1


@pytest.fixture(scope="module")
def t2g_full():
    wl = synth.workload("t2g_k20")
    return synth.build_inputs(wl)


def test_gpu_full_size_against_oracle(t2g_full):
    """3 orbitals, 20^3 k, N = 1080: two points at omega = 0 (one at Gamma, so the local EOM runs too)
    and one at omega = 3, both precision modes, against the numpy oracle."""
    import adga_oracle
    api = _api()
    inp = t2g_full
    hp = adga_oracle.HotPath(inp)
    pts = [(0, 0), (0, 1234), (3, 777)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    ref = hp.run(pts, cl)
    for precision, bubble in (("fp64", 0), ("refined", 0), ("refined", 1)):
        print(f"full size {precision} bubble={bubble}", flush=True)
        with api.Context(inp, bubble=bubble) as ctx:
            out = api.run_loop(ctx, pts, cl, precision)
        for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
            e = rel_err(out[k], ref[k])
            assert e < 1e-10, (precision, k, e)


def test_gpu_full_size_F_zero(t2g_full):
    """chi_loc = chi0^w at N = 1080: eta = 0, so the non-local chi parts and Sigma_nl vanish."""
    import dataclasses
    api = _api()
    inp = dataclasses.replace(t2g_full, wl=dataclasses.replace(t2g_full.wl, noise=0.0))
    pts = [(1, q) for q in range(0, 8000, 97)]
    with api.Context(inp) as ctx:
        out = api.run_loop(ctx, pts, lambda w, c: synth.chi_loc(inp, w, c), "refined")
    scale = np.abs(out["bubble_nl"]).max()
    assert np.abs(out["chi_d"]).max() < 1e-12 * scale
    assert np.abs(out["chi_m"]).max() < 1e-12 * scale
    assert np.abs(out["sigma_nl"]).max() < 1e-14


def test_gpu_refined_equals_fp64_on_a_batch(t2g_full):
    """A 96-point batch at omega = -2 in both modes: per-quantity agreement within 1e-10."""
    api = _api()
    inp = t2g_full
    pts = [(-2, q) for q in range(0, 8000, 83)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    res = {}
    for precision in ("fp64", "refined"):
        with api.Context(inp) as ctx:
            res[precision] = api.run_loop(ctx, pts, cl, precision)
    for k in ("chi_d", "chi_m", "sigma_nl"):
        assert rel_err(res["refined"][k], res["fp64"][k]) < 1e-10, k


def test_gpu_complex64_factors_backward_error(t2g_full):
    """The complex64 LU alone (its Schur updates by the 3M product on 128x128, 128x64, 64x128 and
    32x128 tiles, N = 1080): with a tolerance the first residual already meets, refinement stops
    before any correction, so the reported residual max|B - A y| / max|B| is the backward error of
    the complex64 factors' solve.  Single precision bounds it near c * N * 6e-8; a wrong tile or a
    dropped term of the 3M combination leaves O(1)."""
    api = _api()
    inp = t2g_full
    pts = [(1, q) for q in range(0, 8000, 83)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    with api.Context(inp, refine_tol=0.5) as ctx:
        out = api.run_loop(ctx, pts, cl, "refined")
        st = ctx.stats()
    assert (out["status"] == 0).all()
    assert st["refine_iters"] == 0
    print("complex64 backward error", st["refine_resid"])
    assert 1e-9 < st["refine_resid"] < 2e-5
    with api.Context(inp) as ctx:
        ref = api.run_loop(ctx, pts, cl, "fp64")
    # one complex64 solve, no correction: chi within single-precision reach of the FP64 path
    assert rel_err(out["chi_d"], ref["chi_d"]) < 1e-4


# (workload, vertex noise, overrides): noise 30 (30000x the default) gives condition numbers
# 1e3-1e4 at N = 108 and 200 with most rows swapped; at N = 1620 the same noise reaches 1e7, where
# any two FP64 formulations differ by up to cond * eps ~ 1e-9, so that case uses noise 0.3
# (condition ~ 4e2, still ~400 row swaps).
PIVOT_CASES = [("t2g_k4", 30.0, {}), ("d5_k4", 30.0, {}), ("t2g_k4", 0.3, {"iwfmax_small": 90})]


@pytest.mark.parametrize("name,noise,over", PIVOT_CASES, ids=["t2g_N108", "d5_N200", "t2g_N1620"])
def test_gpu_strong_vertex_pivoting_against_oracle(name, noise, over):
    """A strong vertex makes A = 1 - S C far from the identity (partial pivoting swaps most rows):
    both LU precisions' pivoting paths (complex64 leaves of 8 columns, complex128 of 4; 256- and,
    at N = 1620, 512-thread panels) and several refinement iterations, against the oracle.  The
    refined mode runs with refine_tol = 1e-14 so that its forward error stays below 1e-10 at
    this conditioning."""
    import adga_oracle
    api = _api()
    wl = synth.workload(name, noise=noise, **over)
    inp = synth.build_inputs(wl)
    pts = [(1, 0), (1, 3), (1, 17)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    ref = adga_oracle.HotPath(inp).run(pts, cl)
    for precision in ("fp64", "refined"):
        print(f"strong vertex {name} N={wl.n ** 2 * wl.nf} {precision}", flush=True)
        with api.Context(inp, refine_tol=1e-14) as ctx:
            out = api.run_loop(ctx, pts, cl, precision)
            st = ctx.stats()
        if precision == "refined":
            assert st["refine_iters"] >= 2
        for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
            e = rel_err(out[k], ref[k])
            print(f"  {k} {e:.2e}", flush=True)
            assert e < 1e-10, (precision, k, e)


def test_gpu_multi_stream_lu_is_identical(t2g_full):
    """lu_streams = 2, 3, 4 (the batch's LU split into that many parts on as many streams, staggered by
    one operation, the large Schur updates on the whole batch) run the same kernels on the same data:
    results equal the one-stream run exactly, in both modes, on a 98-point batch (196 matrices: parts
    of 98, 65/65/66 and 49 matrices)."""
    api = _api()
    inp = t2g_full
    pts = [(1, q) for q in range(5, 8000, 82)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    for precision in ("fp64", "refined"):
        res = {}
        for ls in (1, 2, 3, 4):
            with api.Context(inp, lu_streams=ls) as ctx:
                res[ls] = api.run_loop(ctx, pts, cl, precision)
        for ls in (2, 3, 4):
            for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl"):
                assert np.array_equal(res[1][k], res[ls][k]), (precision, ls, k)


def test_gpu_tuned_lanes_are_identical(t2g_full):
    """The default lane count (lu_streams 0, tuned per device and batch shape for the whole process):
    98 points in batches of 16 make six factorisations of 32 matrices and one of 4 per pass (in a fresh
    process the first three of 32 matrices are the timed trials with 1, 2 and 3 lanes, the rest use the
    fastest); every result equals the one-stream run's exactly, in both modes, and an invalid option
    value is refused."""
    api = _api()
    inp = t2g_full
    pts = [(1, q) for q in range(5, 8000, 82)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    for precision in ("fp64", "refined"):
        with api.Context(inp, batch=16, lu_streams=1) as ctx:
            one = api.run_loop(ctx, pts, cl, precision)
        with api.Context(inp, batch=16) as ctx:
            tuned = api.run_loop(ctx, pts, cl, precision)
            ctx.reset_sigma()
            again = api.run_loop(ctx, pts, cl, precision)  # every shape tuned by now
        for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl"):
            assert np.array_equal(one[k], tuned[k]), (precision, k)
            assert np.array_equal(one[k], again[k]), (precision, k)
    with api.Context(inp) as ctx:
        assert api.lib.adga_set_option(ctx.h, b"lu_streams", 5.0) == 1  # ADGA_E_ARG
        assert api.lib.adga_set_option(ctx.h, b"lu_streams", 0.0) == 0


def test_gpu_five_orbital_large_box_against_oracle():
    """5-orbital d-shell model (random symmetric hopping), 16^3 k, 2x80 fermionic frequencies: BSE
    N = 4000, the largest size the batched LU takes (16 rows per thread in the panels; the solves
    run in groups of 2 (complex64) or 1 (complex128) right-hand-side columns).  Two points at two
    frequencies in both precision modes against the numpy oracle."""
    import adga_oracle
    api = _api()
    inp = synth.build_inputs(synth.workload("d5_k16"))
    pts = [(0, 0), (2, 1000)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    print("oracle d5_k16 (N = 4000)", flush=True)
    ref = adga_oracle.HotPath(inp).run(pts, cl)
    for precision in ("fp64", "refined"):
        print(f"d5_k16 {precision}", flush=True)
        with api.Context(inp) as ctx:
            out = api.run_loop(ctx, pts, cl, precision)
        assert (out["status"] == 0).all()
        for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
            e = rel_err(out[k], ref[k])
            assert e < 1e-10, (precision, k, e)


def test_gpu_solve_column_groups_on_a_full_batch():
    """64 points at N = 2000 (5 orbitals, 16^3 k, box 2x40) in one batch: 128 matrices, so the solves'
    n^2 = 25 right-hand-side columns run in column groups on the whole LDS (complex64: 3 groups of 9,
    complex128: 7 of 4), each matrix's groups on one XCD.  Two of the points against the oracle, every
    point of the refined mode against the FP64 mode, and the FP64 mode's own residual."""
    import adga_oracle
    api = _api()
    wl = synth.workload("d5_k16", iwfmax_small=40)
    inp = synth.build_inputs(wl)
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    pts = [(0, q) for q in range(64)]
    check = [0, 37]
    print("oracle d5_k16 box 40 (N = 2000), 2 points", flush=True)
    ref = adga_oracle.HotPath(inp).run([pts[i] for i in check], cl)
    out = {}
    for precision in ("fp64", "refined"):
        print(f"GPU {precision}, 64 points in one batch", flush=True)
        with api.Context(inp, batch=64) as ctx:
            out[precision] = api.run_loop(ctx, pts, cl, precision)
            st = ctx.stats()
        assert (out[precision]["status"] == 0).all()
        assert st["refine_resid"] <= 1e-12, st["refine_resid"]
        for k in ("chi_d", "chi_m", "bubble_nl"):
            e = rel_err(out[precision][k][:, :, check], ref[k])
            assert e < 1e-10, (precision, k, e)
    for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl"):
        assert rel_err(out["refined"][k], out["fp64"][k]) < 1e-10, k


def test_gpu_refinement_hands_unconverged_matrices_to_fp64():
    """refine_maxit = 1 with a tolerance one complex64 correction cannot reach: every matrix is solved
    again in complex128, so the refined mode returns the plain FP64 path's results (the same kernels on
    the same S and C) and counts two re-solved matrices per point."""
    api = _api()
    inp = synth.build_inputs(synth.workload("t2g_k4"))
    pts = [(1, q) for q in range(0, 64, 5)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    with api.Context(inp, refine_maxit=1, refine_tol=1e-17) as ctx:
        out = api.run_loop(ctx, pts, cl, "refined")
        st = ctx.stats()
    with api.Context(inp) as ctx:
        ref = api.run_loop(ctx, pts, cl, "fp64")
    assert st["n_fp64_resolved"] == 2 * len(pts)
    assert st["n_failed_points"] == 0 and (out["status"] == 0).all()
    for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl"):
        assert rel_err(out[k], ref[k]) < 1e-13, k


def test_gpu_near_instability_falls_back_to_fp64():
    """A vertex 1e5 times the default noise at N = 1620 puts A near singular (2-norm condition 3.8e8 for the
    density matrix of q = Gamma, 8e6-2.5e7 elsewhere, from numpy's cond on the oracle's A).  The complex64
    factors still precondition these matrices (the refinement's contraction follows the componentwise
    perturbation, not the normwise condition), but more slowly: the test first runs the refined mode with the
    default limits and requires more than one correction, then again with the iteration limit one below what
    the slowest matrix needed, so that matrix is handed to the complex128 re-solve.  Every point succeeds in
    all runs; the refined results agree with the plain FP64 path, and both with the oracle's explicit LAPACK
    inverse within the conditioning bound cond * eps ~ 1e-7 that any two FP64 formulations share (1e-10 is
    out of reach for any method at this condition number)."""
    import adga_oracle
    api = _api()
    wl = synth.workload("t2g_k4", noise=100.0, iwfmax_small=90)
    inp = synth.build_inputs(wl)
    pts = [(1, 0), (1, 3), (1, 17)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    print("oracle N = 1620, noise 100", flush=True)
    ref = adga_oracle.HotPath(inp).run(pts, cl)
    res, st = {}, {}
    for name, precision, maxit in (("fp64", "fp64", None), ("refined", "refined", None), ("limited", "refined", 0)):
        if maxit == 0:
            assert st["refined"]["refine_iters"] >= 2
            maxit = int(st["refined"]["refine_iters"]) - 1
        print(f"near instability {name} refine_maxit={maxit}", flush=True)
        with api.Context(inp, refine_tol=1e-14, refine_maxit=maxit) as ctx:
            res[name] = api.run_loop(ctx, pts, cl, precision)
            st[name] = ctx.stats()
        print(f"  corrections {st[name]['refine_iters']}, re-solved in FP64 {st[name]['n_fp64_resolved']}", flush=True)
        assert (res[name]["status"] == 0).all()
    assert st["limited"]["n_fp64_resolved"] >= 2  # both matrices of the slowest point
    for k in ("chi_d", "chi_m", "sigma_nl"):
        e_or = rel_err(res["fp64"][k], ref[k])
        e_rf = rel_err(res["refined"][k], res["fp64"][k])
        e_lim = rel_err(res["limited"][k], res["fp64"][k])
        print(f"  {k}: fp64 vs oracle {e_or:.1e}, refined vs fp64 {e_rf:.1e}, limited vs fp64 {e_lim:.1e}", flush=True)
        assert e_or < 1e-6 and e_rf < 1e-6 and e_lim < 1e-6, k


def test_gpu_repeated_q_in_a_batch():
    """A batch whose q list repeats points (q = 5, 9, 5, 17, 9 at omega = 0 and 2): every repeat counts, as
    in the reference's loop over the list.  The k-grid EOM (its scatter then adds the points in order) and
    the direct EOM sums give the same Sigma_nl, and chi of a repeated point equals its first evaluation."""
    api = _api()
    wl = synth.workload("t2g_k4")
    inp = synth.build_inputs(wl)
    pts = [(w, q) for w in (0, 2) for q in (5, 9, 5, 17, 9)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    res = {}
    for eom in (1, 0):
        with api.Context(inp, batch=64, eom=eom) as ctx:
            res[eom] = api.run_loop(ctx, pts, cl, "fp64")
    assert rel_err(res[1]["sigma_nl"], res[0]["sigma_nl"]) < 1e-12
    for k in ("chi_d", "chi_m", "bubble_nl"):
        assert np.array_equal(res[1][k][:, :, 2], res[1][k][:, :, 0]), k
        assert np.array_equal(res[1][k][:, :, 4], res[1][k][:, :, 1]), k


def test_gpu_anisotropic_grid_with_a_long_axis():
    """A 28 x 3 x 2 k grid (q on the same grid) with a stronger vertex (noise 0.05): the k-grid
    transforms of the bubble and EOM run their x passes through the per-output kernel (28 points, beyond
    the per-line kernel's largest line) and the y and z passes per line; chi, bubble_nl and the
    self-energies against the oracle in both precisions, and against the direct sums."""
    import adga_oracle
    api = _api()
    wl = synth.workload("sc1_k8", nk=(28, 3, 2), noise=0.05)
    inp = synth.build_inputs(wl)
    pts = [(0, 0), (0, 37), (0, 90), (2, 101), (-3, 167), (-3, 5)]
    cl = lambda w, c: synth.chi_loc(inp, w, c)  # noqa: E731
    print("anisotropic grid 28 x 3 x 2", flush=True)
    ref = adga_oracle.HotPath(inp).run(pts, cl)
    for precision, modes in (("fp64", (1, 1)), ("refined", (1, 1)), ("fp64", (0, 0))):
        with api.Context(inp, bubble=modes[0], eom=modes[1]) as ctx:
            out = api.run_loop(ctx, pts, cl, precision)
        assert (out["status"] == 0).all()
        for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft"):
            e = rel_err(out[k], ref[k])
            assert e < 1e-10, (precision, modes, k, e)
