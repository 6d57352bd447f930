"""The Fortran host layer (adga_amd/fortran: ISO_C_BINDING modules adga_c, adga_gpu and the drop-in
lapack_module) drives the GPU path and reproduces the reference's golden vectors; its inverse_matrix
keeps the reference's error contract.  The check program is built by `python -m adga_amd.build`."""
import os

import pytest

from conftest import ROOT, load_golden, rel_err
from adga_amd import synth

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "adga_amd", "fortran", "adga_fortran_check")


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["t2g_k4", "sc1_k8"])
def test_fortran_host_matches_reference(name, precision):
    import ref_runner
    from golden.make_golden import FIXTURES
    assert os.path.exists(EXE), "adga_fortran_check not built (python -m adga_amd.build with amdflang)"
    gold = load_golden(name)
    wl = synth.workload(FIXTURES[name][0])
    inp = synth.build_inputs(wl)
    pts = [tuple(p) for p in gold["points"].tolist()]
    print(f"fortran host {name} precision={precision}", flush=True)
    out = ref_runner.run(inp, pts, lambda w, c: synth.chi_loc(inp, w, c), exe=EXE, args=(str(precision),))
    for k in ("chi_d", "chi_m", "bubble_nl", "sigma_nl", "sigma_dmft", "chi_loc_d", "chi_loc_m", "bubble_loc"):
        assert rel_err(out[k], gold[k]) < 1e-10, k
