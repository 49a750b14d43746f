"""ctypes declarations of the translated applications' extern "C" entry points."""
import ctypes

_i, _d, _p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_char_p)

SIGNATURES = {
    "laplace2d": {
        "laplace2d_init": (_i, [_i, _pp, _i, _i, _i]),
        "laplace2d_iterate": (_i, [_i, _p]),
        "laplace2d_fetch_full": (_i, [_p]),
        "laplace2d_finalize": (_i, []),
    },
    "cloverleaf": {
        "cloverleaf_init": (_i, [_i, _pp, ctypes.c_char_p]),
        "cloverleaf_step": (_i, [_i]),
        "cloverleaf_set_graph": (_i, [_i]),
        "cloverleaf_set_chain": (_i, [_i]),
        "cloverleaf_capture": (_i, []),
        "cloverleaf_run": (_i, [_i]),
        "cloverleaf_summary": (_i, [_p]),
        "cloverleaf_state": (_i, [_p]),
        "cloverleaf_fetch": (ctypes.c_long, [ctypes.c_char_p, _p]),
        "cloverleaf_finalize": (_i, []),
    },
    "cloverleaf3d": {
        "cloverleaf3d_init": (_i, [_i, _pp, ctypes.c_char_p]),
        "cloverleaf3d_step": (_i, [_i]),
        "cloverleaf3d_set_graph": (_i, [_i]),
        "cloverleaf3d_capture": (_i, []),
        "cloverleaf3d_run": (_i, [_i]),
        "cloverleaf3d_summary": (_i, [_p]),
        "cloverleaf3d_state": (_i, [_p]),
        "cloverleaf3d_fetch": (ctypes.c_long, [ctypes.c_char_p, _p]),
        "cloverleaf3d_finalize": (_i, []),
    },
    "tealeaf": {
        "tealeaf_init": (_i, [_i, _pp, ctypes.c_char_p]),
        "tealeaf_set_solver": (_i, [_i]),
        "tealeaf_step": (_i, [_i]),
        "tealeaf_run": (_i, [_i]),
        "tealeaf_summary": (_i, [_p]),
        "tealeaf_state": (_i, [_p]),
        "tealeaf_fetch": (ctypes.c_long, [ctypes.c_char_p, _p]),
        "tealeaf_finalize": (_i, []),
        "tealeaf_set_guess": (_i, [_i, _i]),
        "tealeaf_capture_next": (_i, [_i]),
        "tealeaf_guess_state": (_i, [_p]),
        "tealeaf_set_start": (_i, [_p, ctypes.c_long]),
        "tea_unet_init": (_i, [ctypes.c_ulonglong]),
        "tea_unet_num_params": (ctypes.c_long, []),
        "tea_unet_get_params": (_i, [_p]),
        "tea_unet_set_params": (_i, [_p]),
        "tea_unet_forward_host": (_i, [_i, _i, _i, _p, _p]),
        "tea_unet_backward_host": (_i, [_i, _i, _i, _p, _p, _p]),
        "tea_unet_dataset_reserve": (_i, [_i, _i, _i]),
        "tea_unet_dataset_set": (_i, [_i, _p, _p]),
        "tea_unet_train": (_i, [_i, _i, ctypes.c_float, ctypes.c_float, ctypes.c_ulonglong, _p]),
        "tea_unet_free": (_i, []),
    },
    "apitest": {
        "api_mblock": (_i, [_i, _pp, _i, _p, _p]),
        "api_halo3d": (_i, [_i, _pp, _p, _p, _p, _i, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, _p]),
        "api_tridiag": (_i, [_i, _pp, _i, _i, _p, _p, _p, _p, _p, _p, _p, _i, _p]),
        "api_reduce": (_i, [_i, _pp, _i, _i, _p, _p, _p]),
        "api_alias": (_i, [_i, _pp, _i, _i, _p, _p]),
        "api_random": (_i, [_i, _pp, ctypes.c_uint, _i, _i, _p]),
        "api_random_normal": (_i, [_i, _pp, ctypes.c_uint, _i, _i, _p]),
        "api_batch": (_i, [_i, _pp, _i, _i, _p]),
        "api_chain": (_i, [_i, _pp, _i, _i, _i, _i, _p]),
        "api_tridiag2": (_i, [_i, _pp, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p, _p]),
    },
    "multicomp": {
        "multicomp_run": (_i, [_i, _pp, _i, _i, _p]),
    },
    "multicomp_soa": {
        "multicomp_run": (_i, [_i, _pp, _i, _i, _p]),
    },
    "lowdim": {
        "lowdim_run": (_i, [_i, _pp, _i, _i, _i, _p]),
    },
    "mgrid": {
        "mgrid_run": (_i, [_i, _pp, _i, _p, _p, _p, _p, _p]),
    },
    "adi": {
        "adi_init": (_i, [_i, _pp, _i, _p, _i, ctypes.c_ulonglong, _d]),
        "adi_step": (_i, [_i]),
        "adi_fetch": (ctypes.c_long, [_i, _p]),
        "adi_set": (_i, [_i, _p]),
        "adi_finalize": (_i, []),
    },
}


def declare(name, lib):
    for fn, (res, args) in SIGNATURES.get(name, {}).items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
