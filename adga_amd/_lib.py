"""ctypes binding of libadga_amd.so (the C ABI declared in include/adga_amd.h).

The library is built in-tree by `python -m adga_amd.build`.  There is no CPU
fallback: if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libadga_amd.so")

ADGA_OK, ADGA_E_ARG, ADGA_E_HIP, ADGA_E_SINGULAR, ADGA_E_NOCONV, ADGA_E_NONFINITE, ADGA_E_STATE, ADGA_E_NOMEM = range(8)
ADGA_PREC_FP64, ADGA_PREC_REFINED = 0, 1

# every symbol include/adga_amd.h declares
EXPORTS = ["adga_ctx_create", "adga_set_option", "adga_set_local_omega", "adga_run_points", "adga_get_sigma",
           "adga_reset_sigma", "adga_get_stats", "adga_last_error", "adga_ctx_destroy", "adga_inverse_matrix_z",
           "adga_get_giw", "adga_get_nfock", "adga_get_sigma_g_loc", "adga_get_ndga", "adga_set_vq", "adga_set_local_omega_box",
           "adga_set_tangents", "adga_set_local_omega_tangent", "adga_run_points_tangent", "adga_get_sigma_tangent",
           "adga_bubble_q", "adga_chi0_loc", "adga_chi_qw", "adga_eom_dynamic", "adga_eom_dmft", "adga_eom_static",
           "adga_get_gkiw", "adga_get_ndmft", "adga_add_siw_dmft",
           "adga_inverse_matrix_z_noalloc",
           "adga_lambda_create", "adga_lambda_sums", "adga_lambda_target", "adga_lambda_solve", "adga_lambda_chi",
           "adga_lambda_last_error", "adga_lambda_destroy"]


class adga_dims(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nkpx", C.c_int), ("nkpy", C.c_int), ("nkpz", C.c_int),
                ("iwmax", C.c_int), ("iwfmax_small", C.c_int), ("iwbmax_small", C.c_int),
                ("nqp", C.c_int), ("nkp_eom", C.c_int), ("beta", C.c_double), ("mu", C.c_double)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


def _load():
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is not built: run `python -m adga_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(LIBPATH)
    sig = {
        "adga_ctx_create": ([C.POINTER(adga_dims), _ip, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(_vp)], C.c_int),
        "adga_set_option": ([_vp, C.c_char_p, C.c_double], C.c_int),
        "adga_set_vq": ([_vp, _dp], C.c_int),
        "adga_set_tangents": ([_vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "adga_set_local_omega_tangent": ([_vp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "adga_run_points_tangent": ([_vp, C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _ip], C.c_int),
        "adga_get_sigma_tangent": ([_vp, _dp], C.c_int),
        "adga_bubble_q": ([_vp, C.c_int, C.c_int, _dp], C.c_int),
        "adga_chi0_loc": ([_vp, C.c_int, _dp, _dp], C.c_int),
        "adga_chi_qw": ([_vp, _dp, _dp, _dp], C.c_int),
        "adga_eom_dynamic": ([_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "adga_eom_dmft": ([_vp, C.c_int, _dp, _dp], C.c_int),
        "adga_eom_static": ([_vp, C.c_int, _dp, C.c_int, _dp, _dp], C.c_int),
        "adga_get_gkiw": ([_vp, C.c_int, C.c_int, _dp, _dp], C.c_int),
        "adga_get_ndmft": ([_vp, _dp], C.c_int),
        "adga_add_siw_dmft": ([_vp, _dp], C.c_int),
        "adga_set_local_omega": ([_vp, C.c_int, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "adga_set_local_omega_box": ([_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp], C.c_int),
        "adga_run_points": ([_vp, C.c_int, _ip, C.c_int, _dp, _dp, _dp, _ip], C.c_int),
        "adga_get_sigma": ([_vp, _dp, _dp, _dp], C.c_int),
        "adga_reset_sigma": ([_vp], C.c_int),
        "adga_get_stats": ([_vp, _dp, C.c_int], C.c_int),
        "adga_last_error": ([_vp, C.c_char_p, C.c_int], C.c_int),
        "adga_ctx_destroy": ([_vp], C.c_int),
        "adga_inverse_matrix_z": ([C.c_int, _dp, _ip], C.c_int),
        "adga_inverse_matrix_z_noalloc": ([C.c_int, _dp, _ip], C.c_int),
        "adga_get_giw": ([_vp, _dp], C.c_int),
        "adga_get_nfock": ([_vp, _dp], C.c_int),
        "adga_get_sigma_g_loc": ([_vp, _dp, _dp, _dp], C.c_int),
        "adga_get_ndga": ([_vp, _dp, _dp, _dp], C.c_int),
        "adga_lambda_create": ([C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, C.POINTER(_vp)], C.c_int),
        "adga_lambda_sums": ([_vp, C.c_int, _dp, _dp, _dp], C.c_int),
        "adga_lambda_target": ([_vp, _dp], C.c_int),
        "adga_lambda_solve": ([_vp, C.c_double, C.c_double, C.c_int, _dp, _ip], C.c_int),
        "adga_lambda_chi": ([_vp, C.c_double, _dp], C.c_int),
        "adga_lambda_last_error": ([_vp, C.c_char_p, C.c_int], C.c_int),
        "adga_lambda_destroy": ([_vp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


lib = _load()
