"""ctypes binding of libbcg.so (the C ABI of include/bcg.h).

The library is loaded from this package directory only; there is no
fallback: if it is missing or does not load, every call raises.
"""
from __future__ import annotations

import ctypes
import os

from .cells import Params

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbcg.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

ERRORS = {
    -1: "BCG_ERR_ARG",
    -2: "BCG_ERR_CAPACITY",
    -3: "BCG_ERR_OUT_OF_DOMAIN",
    -4: "BCG_ERR_NEIGHBOUR_OVERFLOW",
    -5: "BCG_ERR_HIP",
    -6: "BCG_ERR_UNSUPPORTED",
    -7: "BCG_ERR_STATE",
}

# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "bcg_default_params": (ctypes.c_int, [ctypes.POINTER(Params)]),
    "bcg_create": (ctypes.c_int, [ctypes.POINTER(Params), ctypes.POINTER(_vp)]),
    "bcg_destroy": (ctypes.c_int, [_vp]),
    "bcg_last_error": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int]),
    "bcg_set_cells": (ctypes.c_int, [_vp, _vp, ctypes.c_int32]),
    "bcg_get_cells": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _ip]),
    "bcg_num_cells": (ctypes.c_int, [_vp, _ip]),
    "bcg_set_fields": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "bcg_srand": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "bcg_rand": (ctypes.c_int, [_vp, _ip]),
    "bcg_rand_stream": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, _ip]),
    "bcg_rebuild_neighbours": (ctypes.c_int, [_vp, _ip, ctypes.c_int32]),
    "bcg_get_neighbours": (ctypes.c_int, [_vp, _ip, _ip, ctypes.c_int32]),
    "bcg_sum_forces": (ctypes.c_int, [_vp, _dp, _dp]),
    "bcg_step": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32, _ip, _ip, ctypes.c_int32]),
    "bcg_last_step_stats": (ctypes.c_int, [_vp, _i64p, _ip, _ip]),
    "bcg_last_step_times": (ctypes.c_int, [_vp, _fp]),
}

_lib = None


class BcgError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        self.code = code
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")


def lib():
    """Load libbcg.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m baccolonygrowth_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(SIGNATURES)
