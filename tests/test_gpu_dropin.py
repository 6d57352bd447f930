"""The replacement Fortran modules (adga_amd/fortran: lapack_module; dropin/: one_particle_quant_module,
susc_module, eom_module) are drop-ins: the oracle's loop driver (oracle/ref_harness.f90, which restates
src/main.F90:486-799 and calls the reference's module procedures exactly as main.F90 does, with the
inline zgemm calls on the host) is compiled unchanged against them instead of the reference's modules
(oracle/Makefile target adga_dropin_harness, with the reference's parameters_module, kq_tools and
mpi_org), and reproduces the golden vectors the reference's own modules wrote.  Every replaced procedure
runs on the GPU through the C ABI: accumulate_chi0 (the whole k sum at ik == 1), get_chi0_loc(_inv),
inverse_matrix, calc_chi_qw, calc_eom_static / _dmft / _dynamic, get_nfock."""
import os

import pytest

from conftest import ROOT, load_golden, rel_err
from adga_amd import synth

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "oracle", "_ref", "adga_dropin_harness")

CASES = ["sc1_k8", "two_k4", "t2g_k4", "d5_k4", "t2g_k4_keom", "t2g_k6_qpath", "two_k4_vq", "t2g_k4_vq", "t2g_k4_bub",
         "two_k4_bub"]


@pytest.mark.parametrize("name", CASES)
def test_dropin_modules_reproduce_reference(name):
    import ref_runner
    from golden.make_golden import FIXTURES
    if not os.path.exists(EXE):
        pytest.skip("oracle/_ref/adga_dropin_harness is built only where the reference sources are (make -C oracle)")
    gold = load_golden(name)
    wl = synth.workload(FIXTURES[name][0])
    inp = synth.build_inputs(wl)
    pts = [tuple(p) for p in gold["points"].tolist()]
    print(f"drop-in modules {name}: {len(pts)} points", flush=True)
    out = ref_runner.run(inp, pts, lambda w, c: synth.chi_loc(inp, w, c), exe=EXE, timeout=600)
    keys = ["chi_d", "chi_m", "bubble_nl", "chi_loc_d", "chi_loc_m", "bubble_loc", "gammaqd0", "eta_d0", "eta_m0"]
    if wl.do_eom:
        keys += ["sigma_nl", "sigma_dmft"]
        if inp.v_q is not None:
            keys.append("sigma_hf")
    for k in keys:
        e = rel_err(out[k], gold[k])
        assert e < 1e-10, (k, e)
