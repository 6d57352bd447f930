"""adga_inverse_matrix_z_noalloc (no device allocation: its buffers live in a file-scope device array)
against adga_inverse_matrix_z, both through the C ABI, on two calls in a row, the second with a smaller
matrix (the array keeps the first call's contents), plus the size limit."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

NOALLOC_MAX_N = 1024  # ADGA_INVERSE_NOALLOC_MAX_N


def _noalloc():
    from adga_amd import _lib
    f = _lib.lib.adga_inverse_matrix_z_noalloc
    f.argtypes = [C.c_int, _lib._dp, _lib._ip]
    f.restype = C.c_int
    return f, _lib


def _call(f, _lib, a):
    m = np.array(a, dtype=np.complex128, order="F", copy=True)
    ierr = C.c_int(-1)
    rc = f(m.shape[0], m.ctypes.data_as(_lib._dp), C.byref(ierr))
    return rc, ierr.value, m


def test_noalloc_inverse_matches_allocating_entry_point():
    from adga_amd import api
    f, _lib = _noalloc()
    rng = np.random.default_rng(11)
    for n in (NOALLOC_MAX_N, 300):  # the second call is smaller than the first
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)) + 2 * np.eye(n)
        print(f"noalloc inverse n = {n}", flush=True)
        rc, ierr, inv = _call(f, _lib, a)
        assert rc == 0 and ierr == 0
        ref, ierr_ref = api.inverse_matrix(a)
        assert ierr_ref == 0
        assert rel_err(inv, ref) < 1e-13
        assert rel_err(inv, np.linalg.inv(a)) < 1e-11


def test_noalloc_inverse_rejects_large_and_singular():
    f, _lib = _noalloc()
    rc, _, _ = _call(f, _lib, np.eye(NOALLOC_MAX_N + 1))
    assert rc == _lib.ADGA_E_ARG
    rc, ierr, _ = _call(f, _lib, np.ones((8, 8)))
    assert rc == _lib.ADGA_E_SINGULAR and ierr > 0
    # the array keeps the singular call's contents: a good call afterwards is still exact
    a = np.eye(5) * (2.0 + 1.0j)
    rc, ierr, inv = _call(f, _lib, a)
    assert rc == 0 and ierr == 0
    assert rel_err(inv, np.linalg.inv(a)) < 1e-15
