"""The C-ABI library loads and exports every entry point include/adga_amd.h declares; argument
checks that need no GPU return ADGA_E_ARG."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "adga_amd.h")).read()
    return sorted(set(re.findall(r"^int\s+(adga_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "adga_run_points" in names and "adga_inverse_matrix_z" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    from adga_amd import _lib
    lib = C.CDLL(_lib.LIBPATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTS) == _declared()


def test_fortran_layer_binds_every_declared_entry_point():
    """adga_amd/fortran/adga_c.f90 has a bind(c) interface for each entry point of the header, so a
    Fortran host reaches all of them."""
    src = open(os.path.join(ROOT, "adga_amd", "fortran", "adga_c.f90")).read()
    bound = sorted(set(re.findall(r"bind\(c,\s*name='(adga_\w+)'\)", src)))
    assert bound == _declared()


def test_argument_validation_without_gpu():
    from adga_amd import _lib
    lib = _lib.lib
    d = _lib.adga_dims(0, 4, 4, 4, 10, 2, 1, 64, 64, 10.0, 0.0)  # ndim = 0 is invalid
    h = C.c_void_p()
    z = (C.c_double * 8)()
    i = (C.c_int * 8)(1, 1, 1, 1, 1, 1, 1, 1)
    rc = lib.adga_ctx_create(C.byref(d), i, i, z, z, z, z, z, z, z, C.byref(h))
    assert rc == _lib.ADGA_E_ARG and not h.value
    d = _lib.adga_dims(3, 4, 4, 4, 2, 2, 1, 64, 64, 10.0, 0.0)  # iwmax < iwfmax_small + iwbmax_small
    assert lib.adga_ctx_create(C.byref(d), i, i, z, z, z, z, z, z, z, C.byref(h)) == _lib.ADGA_E_ARG
    ierr = C.c_int(0)
    assert lib.adga_inverse_matrix_z(0, z, C.byref(ierr)) == _lib.ADGA_E_ARG
    assert lib.adga_inverse_matrix_z(4097, z, C.byref(ierr)) == _lib.ADGA_E_ARG  # above ADGA_INVERSE_MAX_N
    assert lib.adga_ctx_destroy(None) == _lib.ADGA_OK


def test_hdf5_library_exports_every_declared_symbol():
    h5io = pytest.importorskip("adga_amd.h5io", reason="libadga_h5.so not built (needs the HDF5 C library)")
    src = open(os.path.join(ROOT, "include", "adga_h5.h")).read()
    declared = sorted(set(re.findall(r"^int\s+(adga_h5_\w+)\s*\(", src, flags=re.M)))
    lib = C.CDLL(h5io.LIBPATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(h5io.EXPORTS) == declared
