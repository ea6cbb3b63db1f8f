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
    "bcg_get_fields": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _dp]),
    "bcg_srand": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "bcg_rand": (ctypes.c_int, [_vp, _ip]),
    "bcg_rand_stream": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, _ip]),
    "bcg_rebuild_neighbours": (ctypes.c_int, [_vp, _ip, ctypes.c_int32]),
    "bcg_get_neighbours": (ctypes.c_int, [_vp, _ip, _ip, ctypes.c_int32]),
    "bcg_sum_forces": (ctypes.c_int, [_vp, _dp, _dp]),
    "bcg_mean_stress": (ctypes.c_int, [_vp, _dp, _dp]),
    "bcg_step": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32, _ip, _ip, ctypes.c_int32]),
    "bcg_last_step_stats": (ctypes.c_int, [_vp, _i64p, _ip, _ip]),
    "bcg_step_out": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32, _vp, ctypes.c_int32, _ip,
                                    _ip, ctypes.c_int32]),
    "bcg_last_step_times": (ctypes.c_int, [_vp, _fp]),
}


class RD3DParams(ctypes.Structure):
    """Mirror of `bcg_rd3d_params` (include/bcg.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("h", ctypes.c_double),
        ("origin", ctypes.c_double * 3),
        ("D", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("K", ctypes.c_double),
        ("c_res", ctypes.c_double),
        ("nu1", ctypes.c_int32),
        ("nu2", ctypes.c_int32),
        ("coarse_sweeps", ctypes.c_int32),
        ("max_levels", ctypes.c_int32),
        ("inner", ctypes.c_int32),
        ("smoother", ctypes.c_int32),
    ]


SIGNATURES.update(
    {
        "bcg_rd3d_default_params": (ctypes.c_int, [ctypes.POINTER(RD3DParams)]),
        "bcg_rd3d_create": (ctypes.c_int, [ctypes.POINTER(RD3DParams), ctypes.POINTER(_vp)]),
        "bcg_rd3d_destroy": (ctypes.c_int, [_vp]),
        "bcg_rd3d_last_error": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int]),
        "bcg_rd3d_levels": (ctypes.c_int, [_vp, _ip, _ip, ctypes.c_int32]),
        "bcg_rd3d_set_field": (ctypes.c_int, [_vp, _dp]),
        "bcg_rd3d_get_field": (ctypes.c_int, [_vp, _dp]),
        "bcg_rd3d_set_uptake": (ctypes.c_int, [_vp, _dp]),
        "bcg_rd3d_get_uptake": (ctypes.c_int, [_vp, _dp]),
        "bcg_rd3d_deposit": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int32, _ip]),
        "bcg_rd3d_step": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _ip, _dp]),
        "bcg_rd3d_vcycle": (ctypes.c_int, [_vp, ctypes.c_int32, _fp]),
        "bcg_rd3d_last_times": (ctypes.c_int, [_vp, _fp]),
        "bcg_rd3d_set_state": (ctypes.c_int, [_vp, _dp, _dp]),
        "bcg_rd3d_set_coeffs": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double]),
        "bcg_rd3d_deposit_cells": (ctypes.c_int, [_vp, _vp, ctypes.c_double, _ip]),
        "bcg_rd3d_growth_to_cells": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double]),
        "bcg_rd3d_update_cells": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                 ctypes.c_int32]),
        "bcg_rd3d_adjoint": (
            ctypes.c_int,
            [_vp, _dp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, _dp, _dp, _dp, _dp],
        ),
        "bcg_rd3d_gather": (ctypes.c_int, [_vp, _dp, ctypes.c_int32, _dp, _dp]),
        "bcg_rd3d_solve": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, ctypes.c_double, ctypes.c_int32, _dp, _ip, _dp]),
        "bcg_rd3d_deposit_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, _vp, _vp]),
    }
)

ENV_CELL_FIELDS = [
    "Carbon", "Waste", "GrowthRate1", "GrowthRate2", "Oxygen", "Acetate",
    "pH", "Buffer", "Acetic", "Hydrox",
    "qCarmnt", "qAcemnt", "pAcemnt", "qOxymnt",
    "GrCarAer", "GrCarAna", "GrAceAer",
]
AGA_FIELDS = ["CarbonAgar", "WasteAgar", "OxygenAgar", "AcetateAgar"]


class EnvParams(ctypes.Structure):
    """Mirror of `bcg_env_params` (include/bcg.h)."""

    _fields_ = [
        (n, ctypes.c_int32)
        for n in (
            "BoxX", "BoxY", "BoxZ", "BoxZAgar", "maxLevelsMG", "refinementGridHeight", "FilterLen",
            "crossFeeding", "NutrientGSI", "InterfaceCondition",
        )
    ] + [
        (n, ctypes.c_double)
        for n in (
            "BoxLength", "cellRadius", "density_threshold", "UpdateTime", "Cdt",
            "KC_aer", "KC_ana", "KW", "KO", "KA", "KI",
            "C_rate", "W_rate", "Carbon_to_Waste", "Maintenance_rate",
            "maxGrowthRate1", "maxGrowthRate2",
            "DiffColony", "DiffAgar", "maxCarbon", "maxO2", "maxO2_Agar",
            "qCarAer", "qCarAna", "qAceAer", "qO2Car", "qO2Ace", "pAceAer", "pAceAna",
            "DiffColonyCarbon", "DiffAgarCarbon", "DiffColonyO2", "DiffAgarO2", "DiffColonyAce", "DiffAgarAce",
            "maxGrowthRateCarAer", "maxGrowthRateCarAna", "maxGrowthRateAceAer",
            "Kw", "Ka", "Kb", "Cb", "c1",
            "qCarAernt", "qCarAnant", "qAceAernt", "qO2Carnt", "qO2Acent", "pAceAnant",
        )
    ]


class Clock(ctypes.Structure):
    """Mirror of `bcg_clock` (include/bcg.h)."""

    _fields_ = [
        (n, ctypes.c_double)
        for n in (
            "t", "dt", "OutputTime", "UpdateTime", "HeightDensityUpdateTime",
            "NextOutTime", "NextUpdateTime", "HeightDensityNextUpdateTime",
        )
    ] + [(n, ctypes.c_int32) for n in ("OutFlag", "UpdateFlag", "HeightDensityUpdateFlag")]


SIGNATURES.update(
    {
        "bcg_env_default_params": (ctypes.c_int, [ctypes.POINTER(EnvParams)]),
        "bcg_env_create": (ctypes.c_int, [ctypes.POINTER(EnvParams), _vp, ctypes.POINTER(_vp)]),
        "bcg_env_destroy": (ctypes.c_int, [_vp]),
        "bcg_env_last_error": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int]),
        "bcg_env_set_lds_limit": (ctypes.c_int, [_vp, ctypes.c_int64]),
        "bcg_env_set_execution": (ctypes.c_int, [_vp, ctypes.c_int32]),
        "bcg_env_set_state": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "bcg_env_get_state": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
        "bcg_env_fields": (ctypes.c_int, [_vp, ctypes.c_int32]),
        "bcg_env_nutrients": (ctypes.c_int, [_vp, _ip]),
        "bcg_env_get_fields": (ctypes.c_int, [_vp, _dp, _dp, _dp, _dp, _ip]),
        "bcg_clock_init": (
            ctypes.c_int,
            [ctypes.POINTER(Clock), ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double],
        ),
        "bcg_clock_advance": (ctypes.c_int, [ctypes.POINTER(Clock)]),
        "bcg_sim_step": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Clock), _ip, _ip, ctypes.c_int32]),
        "bcg_sim_prepare": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Clock)]),
        "bcg_sim_advance": (ctypes.c_int, [_vp, ctypes.POINTER(Clock), _vp, ctypes.c_int32, _ip, _ip, ctypes.c_int32]),
    }
)

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
