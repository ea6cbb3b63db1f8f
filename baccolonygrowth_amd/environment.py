"""Planar colony environment on the GPU: density, height and normal fields,
and the nutrient update (the bcg_env_* C ABI), plus the time-loop clock and
one pass of the reference time loop (bcg_clock_*, bcg_sim_step, or
bcg_sim_prepare + bcg_sim_advance around the output step).

Mirrors the reference's field/nutrient seam: GetDensity / ShiftDensity /
BoxAverage / Height_Average / Smooth / GetSurfaceNormal
(ABM/Compute.cpp:120-738) and UpdateEnvArrayMG_OxygenAcetate
(ABM/Multigrid.cpp:984-1125), driven as in ABM/Simulation.cpp:187-312.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import AGA_FIELDS, ENV_CELL_FIELDS, BcgError, Clock, EnvParams, _vp, lib
from .output import ENV_OUTPUT_FIELDS, write_density_text, write_env_text  # noqa: F401 (output-step writers)

ENV_DTYPE = np.dtype([(f, "<f8") for f in ENV_CELL_FIELDS])  # struct LocalEnv, 136 B
AGA_DTYPE = np.dtype([(f, "<f8") for f in AGA_FIELDS])  # struct LocalAga, 32 B

ENV_INIT, ENV_UPDATE, ENV_HEIGHT = 0, 1, 2

# input-file key (ABM/InputOutput.cpp:481-661) -> bcg_env_params field
_KEYS = {
    "Box_x": "BoxX", "Box_y": "BoxY", "Box_z": "BoxZ", "Box_z_agar": "BoxZAgar", "maxLevelsMG": "maxLevelsMG",
    "refinementGridHeight": "refinementGridHeight", "FilterLen": "FilterLen", "CrossFeeding": "crossFeeding",
    "NutrientGSI": "NutrientGSI", "InterfaceCondition": "InterfaceCondition", "Radius": "cellRadius",
    "Density_Threshold": "density_threshold", "Update_Time": "UpdateTime", "Cdt": "Cdt", "KC_aer": "KC_aer",
    "KC_ana": "KC_ana", "KW": "KW", "KO": "KO", "KA": "KA", "KI": "KI", "C_rate": "C_rate", "W_rate": "W_rate",
    "Carbon_to_Waste": "Carbon_to_Waste", "MaintenanceRate": "Maintenance_rate", "Growth_Rate1": "maxGrowthRate1",
    "Growth_Rate2": "maxGrowthRate2", "Diff_Colony": "DiffColony", "Diff_Agar": "DiffAgar", "maxCarbon": "maxCarbon",
    "Max_Oxygen": "maxO2", "Q_Carbon_Aerobic": "qCarAer", "Q_Carbon_Anaerobic": "qCarAna",
    "Q_Acetate_Aerobic": "qAceAer", "Q_Oxygen_Carbon": "qO2Car", "Q_Oxygen_Acetate": "qO2Ace",
    "P_Acetate_Aerobic": "pAceAer", "P_Acetate_Anaerobic": "pAceAna", "Diff_Colony_Carbon": "DiffColonyCarbon",
    "Diff_Agar_Carbon": "DiffAgarCarbon", "Diff_Colony_Oxygen": "DiffColonyO2", "Diff_Agar_Oxygen": "DiffAgarO2",
    "Diff_Colony_Acetate": "DiffColonyAce", "Diff_Agar_Acetate": "DiffAgarAce",
    "Max_Growth_Rate_Carbon_Aerobic": "maxGrowthRateCarAer", "Max_Growth_Rate_Carbon_Anaerobic": "maxGrowthRateCarAna",
    "Max_Growth_Rate_Acetate_Aerobic": "maxGrowthRateAceAer", "Kw": "Kw", "Ka": "Ka", "Kb": "Kb", "Cb": "Cb",
    "c1": "c1", "qCarAernt": "qCarAernt", "qCarAnant": "qCarAnant", "qAceAernt": "qAceAernt",
    "qO2Carnt": "qO2Carnt", "qO2Acent": "qO2Acent", "pAceAnant": "pAceAnant",
}


def default_env_params(**kw) -> EnvParams:
    p = EnvParams()
    rc = lib().bcg_env_default_params(ctypes.byref(p))
    if rc:
        raise BcgError(rc)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# bcg_env_set_execution flags (include/bcg.h)
EXEC_ONE_WORKGROUP = 1  # the whole update in one workgroup
EXEC_NO_GRAPH = 2  # every phase launched directly
EXEC_WG_SWEEPS = 4  # an outer iteration's sweeps in one workgroup
EXEC_GRID_SWEEPS = 8  # one device-wide launch per sweep


def env_params_from_file(path: str) -> EnvParams:
    """bcg_env_params from a reference input file: defaults of
    ABM/Constants.cpp overlaid with the file's keys (atof/atoi semantics of
    ABM/InputOutput.cpp:450-677: a number prefix, trailing text ignored), and
    BoxLength = L_divide + 2 Radius (ABM/CellMD3D.cpp:52)."""
    import re

    p = default_env_params()
    # the reference defaults where the demo file differs from them
    ref_defaults = dict(BoxX=150, BoxY=150, BoxZ=50, BoxZAgar=30, maxLevelsMG=1, refinementGridHeight=4,
                        FilterLen=5, crossFeeding=1, NutrientGSI=0, InterfaceCondition=1, cellRadius=0.34,
                        density_threshold=0.6, UpdateTime=100 * 0.00001, Cdt=0.1, KC_aer=0.001, KC_ana=0.001,
                        KW=0.001, KO=1.0, KA=1.0, KI=1.0, C_rate=3.7e-5 * 2.0 * 5.0, W_rate=3.7e-5 * 2.0 * 5.0,
                        Carbon_to_Waste=1.0, Maintenance_rate=0.0, maxGrowthRate1=1.0, maxGrowthRate2=1.0,
                        DiffColony=1.0, DiffAgar=3.0, maxCarbon=0.01, maxO2=1.0, maxO2_Agar=0.26)
    for k in ("qCarAer", "qCarAna", "qAceAer", "qO2Car", "qO2Ace", "pAceAer", "pAceAna", "DiffColonyCarbon",
              "DiffAgarCarbon", "DiffColonyO2", "DiffAgarO2", "DiffColonyAce", "DiffAgarAce", "maxGrowthRateCarAer",
              "maxGrowthRateCarAna", "maxGrowthRateAceAer", "Kw", "Ka", "Kb", "Cb", "c1", "qCarAernt", "qCarAnant",
              "qAceAernt", "qO2Carnt", "qO2Acent", "pAceAnant"):
        ref_defaults[k] = 1.0
    for k, v in ref_defaults.items():
        setattr(p, k, v)
    L_divide, radius = 5.75, 0.34
    num = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?")
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) < 2:
                continue
            key, val = parts[0], parts[1]
            m = num.match(val)
            x = float(m.group(0)) if m else 0.0
            if key == "L_divide":
                L_divide = x
            if key == "Radius":
                radius = x
            if key == "Box_Dim":
                p.BoxX = p.BoxY = int(x)
            if key in _KEYS:
                f = _KEYS[key]
                setattr(p, f, int(x) if isinstance(getattr(p, f), int) else x)
    p.BoxLength = L_divide + 2 * radius
    return p


class Environment:
    """Environment fields of a Colony (mechanics context) on the GPU."""

    def __init__(self, colony, params: EnvParams):
        self.colony = colony
        self.p = params
        h = _vp()
        rc = lib().bcg_env_create(ctypes.byref(params), colony._h, ctypes.byref(h))
        if rc:
            buf = ctypes.create_string_buffer(512)
            lib().bcg_last_error(colony._h, buf, 512)  # the create error is kept on the mechanics context
            raise BcgError(rc, buf.value.decode())
        self._h = h

    def _check(self, rc, h="self"):
        if rc:
            msg = ""
            hh = self._h if h == "self" else h
            if hh:
                buf = ctypes.create_string_buffer(512)
                lib().bcg_env_last_error(hh, buf, 512)
                msg = buf.value.decode()
            raise BcgError(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            lib().bcg_env_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sizes(self):
        p = self.p
        return p.BoxX * p.BoxZ, p.BoxX * p.BoxZAgar, p.BoxX

    def set_state(self, colony=None, colony_old=None, agar=None, agar_old=None, wall=None, wall_old=None):
        nb, na, nw = self._sizes()
        ml = self.p.maxLevelsMG
        args, keep = [], []
        for a, dt, n in ((colony, ENV_DTYPE, nb), (colony_old, ENV_DTYPE, nb), (agar, AGA_DTYPE, ml * na),
                         (agar_old, AGA_DTYPE, ml * na), (wall, AGA_DTYPE, ml * nw), (wall_old, AGA_DTYPE, ml * nw)):
            if a is None:
                args.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=dt).reshape(-1)
            if a.size != n:
                raise ValueError(f"expected {n} records, got {a.size}")
            keep.append(a)
            args.append(ctypes.c_void_p(a.ctypes.data))
        self._check(lib().bcg_env_set_state(self._h, *args))

    def state(self):
        nb, na, nw = self._sizes()
        ml = self.p.maxLevelsMG
        st = {
            "colony": np.zeros(nb, ENV_DTYPE),
            "colony_old": np.zeros(nb, ENV_DTYPE),
            "agar": np.zeros(ml * na, AGA_DTYPE),
            "agar_old": np.zeros(ml * na, AGA_DTYPE),
            "wall": np.zeros(ml * nw, AGA_DTYPE),
            "wall_old": np.zeros(ml * nw, AGA_DTYPE),
        }
        self._check(lib().bcg_env_get_state(self._h, *[ctypes.c_void_p(a.ctypes.data) for a in st.values()]))
        return st

    def set_lds_limit(self, nbytes: int):
        """Cap the LDS the nutrient update may use for the swept colony boxes
        (0: global memory); the results do not depend on it."""
        self._check(lib().bcg_env_set_lds_limit(self._h, int(nbytes)))

    def set_execution(self, flags: int):
        """EXEC_* flags: how the nutrient update runs; the results do not
        depend on them."""
        self._check(lib().bcg_env_set_execution(self._h, int(flags)))

    def fields(self, pass_: int = ENV_UPDATE):
        self._check(lib().bcg_env_fields(self._h, pass_))

    def nutrients(self) -> int:
        c = ctypes.c_int32(0)
        self._check(lib().bcg_env_nutrients(self._h, ctypes.byref(c)))
        return c.value

    def get_fields(self):
        nb, _, nw = self._sizes()
        out = {k: np.zeros(nb) for k in ("density", "density_shift", "inside")}
        out["wall_shift"] = np.zeros(nw)
        b = np.zeros(5, dtype=np.int32)
        P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
        self._check(lib().bcg_env_get_fields(self._h, P(out["density"]), P(out["density_shift"]), P(out["inside"]),
                                             P(out["wall_shift"]), b.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        out["bounds"] = b
        return out


def clock(t0: float, dt: float, output_time: float, update_time: float) -> Clock:
    c = Clock()
    rc = lib().bcg_clock_init(ctypes.byref(c), t0, dt, output_time, update_time)
    if rc:
        raise BcgError(rc)
    return c


def advance(c: Clock) -> None:
    rc = lib().bcg_clock_advance(ctypes.byref(c))
    if rc:
        raise BcgError(rc)


def _raise(colony, env, rc):
    msgs = []
    for h, fn in ((colony._h, lib().bcg_last_error), (env._h if env is not None else None, lib().bcg_env_last_error)):
        if h:
            buf = ctypes.create_string_buffer(512)
            fn(h, buf, 512)
            if buf.value:
                msgs.append(buf.value.decode(errors="replace"))
    raise BcgError(rc, "; ".join(msgs))


def sim_prepare(colony, env, c: Clock) -> None:
    """The time loop's pass up to its output point (bcg_sim_prepare): the
    UpdateFlag fields, nutrients and list rebuild, the HeightDensity fields."""
    rc = lib().bcg_sim_prepare(colony._h, env._h if env is not None else None, ctypes.byref(c))
    if rc:
        _raise(colony, env, rc)


def sim_advance(colony, c: Clock, moved: bool = False, lineage_cap: int = 4096):
    """The rest of the pass (bcg_sim_advance): move + grow, divisions, clock.
    Returns (lineage pairs, the moved cells before the divisions or None)."""
    from .cells import CELL_DTYPE

    nd = ctypes.c_int32(0)
    lin = np.zeros(2 * lineage_cap, dtype=np.int32)
    mv = np.zeros(colony.n, CELL_DTYPE) if moved else None
    rc = lib().bcg_sim_advance(colony._h, ctypes.byref(c), mv.ctypes.data_as(ctypes.c_void_p) if moved else None,
                               len(mv) if moved else 0, ctypes.byref(nd),
                               lin.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), lineage_cap)
    if rc:
        _raise(colony, None, rc)
    return lin[: 2 * min(nd.value, lineage_cap)].reshape(-1, 2), mv


def sim_step(colony, env, c: Clock, lineage_cap: int = 4096):
    """One pass of the reference time loop (bcg_sim_step); returns the
    (parent, daughter) pairs of this step's divisions."""
    nd = ctypes.c_int32(0)
    lin = np.zeros(2 * lineage_cap, dtype=np.int32)
    rc = lib().bcg_sim_step(colony._h, env._h if env is not None else None, ctypes.byref(c), ctypes.byref(nd),
                            lin.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), lineage_cap)
    if rc:
        _raise(colony, env, rc)
    return lin[: 2 * min(nd.value, lineage_cap)].reshape(-1, 2)
