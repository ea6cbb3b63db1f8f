"""Cell records, parameters and the reference's on-disk cell formats.

`CELL_DTYPE` is byte-identical to the reference's `struct Cell`
(AgentBasedModel-BacColonyGrowth/Cell.h:15-28, with `Segment` from
tools.h:71-87), 208 bytes. The Restart binary (InputOutput.cpp:394-448) is a
16-byte header (t, dt as float64) followed by Cell[N].

`Params` is the ctypes mirror of `bcg_params` in include/bcg.h. Its field
names follow the reference globals (Constants.cpp:1-110); `read_parameters`
parses the reference's `name value` input files with the semantics of
ReadParameters (InputOutput.cpp:450-677): unknown names are reported and
ignored, values go through atof/atoi.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

CELL_DTYPE = np.dtype(
    [
        ("p", "<f8", 3),
        ("q", "<f8", 3),
        ("time_p", "<f8"),
        ("time_q", "<f8"),
        ("age_p", "<i4"),
        ("age_q", "<i4"),
        ("Length", "<f8"),
        ("Ldiv", "<f8"),
        ("Radius", "<f8"),
        ("GrowthRate", "<f8"),
        ("Velocity", "<f8", 3),
        ("AngularVelocity", "<f8", 3),
        ("DynFric", "<f8", 3),
        ("StaFric", "<f8", 3),
        ("Type", "<i4"),
        ("Ancestor", "<i4"),
    ]
)
assert CELL_DTYPE.itemsize == 208


class Params(ctypes.Structure):
    """Mirror of `bcg_params` (include/bcg.h)."""

    _fields_ = [
        ("cellRadius", ctypes.c_double),
        ("L_divide", ctypes.c_double),
        ("k_cc", ctypes.c_double),
        ("k_wc", ctypes.c_double),
        ("varL", ctypes.c_double),
        ("varAngle", ctypes.c_double),
        ("var_pos", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("wall_rough", ctypes.c_double),
        ("gamma_t", ctypes.c_double),
        ("wall_mu", ctypes.c_double),
        ("cell_mu", ctypes.c_double),
        ("density_threshold", ctypes.c_double),
        ("tension", ctypes.c_double),
        ("DH", ctypes.c_double),
        ("Rc", ctypes.c_double),
        ("BoxLength", ctypes.c_double),
        ("BoxX", ctypes.c_int32),
        ("BoxY", ctypes.c_int32),
        ("BoxZ", ctypes.c_int32),
        ("refinementGridHeight", ctypes.c_int32),
        ("maxCells", ctypes.c_int32),
        ("maxNeighbours", ctypes.c_int32),
        ("crossFeeding", ctypes.c_int32),
        ("planar", ctypes.c_int32),
    ]

    def copy(self) -> "Params":
        p = Params()
        ctypes.pointer(p)[0] = self
        return p

    def replace(self, **kw) -> "Params":
        p = self.copy()
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return p

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


# Input-file key -> Params field (the subset of InputOutput.cpp:481-661 that
# the mechanics reads). Keys for the nutrient chemistry are kept in
# `extra` by read_parameters.
_KEYS_F = {
    "Radius": "cellRadius",
    "L_divide": "L_divide",
    "k_cc": "k_cc",
    "k_wc": "k_wc",
    "var_L": "varL",
    "var_angle": "varAngle",
    "var_pos": "var_pos",
    "Viscosity": "viscosity",
    "Wall_Rough": "wall_rough",
    "Gamma": "gamma_t",
    "Wall_Mu": "wall_mu",
    "Cell_Mu": "cell_mu",
    "Density_Threshold": "density_threshold",
    "Surface_Tension": "tension",
    "Delta_H": "DH",
    "Rc": "Rc",
}
_KEYS_I = {
    "Box_x": "BoxX",
    "Box_y": "BoxY",
    "Box_z": "BoxZ",
    "refinementGridHeight": "refinementGridHeight",
    "MaxCells": "maxCells",
    "CrossFeeding": "crossFeeding",
}


def reference_defaults() -> Params:
    """Globals of Constants.cpp:1-110 (before any input file is read)."""
    p = Params()
    p.cellRadius = 0.34
    p.L_divide = 5.75
    p.k_cc = 100000.0
    p.k_wc = 100000.0
    p.varL = 0.0
    p.varAngle = 0.0
    p.var_pos = 0.0
    p.viscosity = 1.0
    p.wall_rough = 0.0
    p.gamma_t = 10000.0
    p.wall_mu = 0.15
    p.cell_mu = 0.15
    p.density_threshold = 0.6
    p.tension = 10.0
    p.DH = 0.0
    p.Rc = 3.0
    p.BoxX, p.BoxY, p.BoxZ = 150, 150, 50
    p.refinementGridHeight = 4
    p.maxCells = 100000
    p.crossFeeding = 1
    p.planar = 1
    p.maxNeighbours = 64
    p.BoxLength = p.L_divide + 2 * p.cellRadius
    return p


def _atoi(s: str) -> int:
    """C atoi: leading integer prefix, 0 if none."""
    s = s.strip()
    n = 0
    sign = 1
    i = 0
    if i < len(s) and s[i] in "+-":
        sign = -1 if s[i] == "-" else 1
        i += 1
    digits = ""
    while i < len(s) and s[i].isdigit():
        digits += s[i]
        i += 1
    if digits:
        n = int(digits)
    return sign * n


def _atof(s: str) -> float:
    """C atof on the longest numeric prefix (a trailing ';' is ignored)."""
    s = s.strip()
    for end in range(len(s), 0, -1):
        try:
            return float(s[:end])
        except ValueError:
            continue
    return 0.0


def read_parameters(path: str, base: Params | None = None):
    """Parse a reference input file (InputOutput.cpp:450-677).

    Returns (params, extra) where `extra` maps every other recognised key to
    its value as the reference parses it. BoxLength is set as the reference
    derives it, L_divide + 2*Radius (CellMD3D.cpp:52).
    """
    p = (base or reference_defaults()).copy()
    extra: dict[str, float] = {}
    with open(path, "r") as fh:
        for line in fh:
            parts = line.split()
            if len(parts) < 2:
                continue
            name, value = parts[0], parts[1]
            if name in _KEYS_F:
                setattr(p, _KEYS_F[name], _atof(value))
            elif name in _KEYS_I:
                v = _atoi(value)
                if name == "CrossFeeding":
                    v = 1 if v else 0
                setattr(p, _KEYS_I[name], v)
            elif name == "Restart_dir":
                extra[name] = value  # strcpy of the token (InputOutput.cpp:642-643)
            elif name == "Box_Dim":
                p.BoxX = p.BoxY = _atoi(value)
            else:
                extra[name] = _atof(value)
    p.BoxLength = p.L_divide + 2 * p.cellRadius
    return p, extra


def max_cells_per_box(p: Params) -> int:
    """Box capacity rule of CellMD3D.cpp:54."""
    return int(
        math.ceil(
            p.BoxLength ** 3
            / (math.pi * p.cellRadius * p.cellRadius * p.L_divide * p.density_threshold / 2)
        )
    )


def load_restart(path: str):
    """Read a Restart binary (InputOutput.cpp:394-431): (t, dt, cells)."""
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < 16 or (raw.size - 16) % CELL_DTYPE.itemsize:
        raise ValueError(f"{path}: not a restart file ({raw.size} bytes)")
    t, dt = np.frombuffer(raw[:16].tobytes(), "<f8")
    cells = np.frombuffer(raw[16:].tobytes(), CELL_DTYPE).copy()
    return float(t), float(dt), cells


# One line of the Cells/<k>.txt output (ABM/InputOutput.cpp:799-804): t, ID,
# Type, p, q, Length, the stress diagonal xx yy zz, Velocity, GrowthRate,
# DynFric, StaFric, time_p, time_q, age_p, age_q, Ancestor.
CELLS_LINE = ("%4.4f %d %d %4.4f %4.4f %4.4f %4.4f %4.4f %4.4f %4.4f %4.4f %4.4E %4.4E %4.4E %4.4E %4.4E %4.4E "
              "%4.4f %4.4f %4.4f %4.4f %4.4f %4.4f %4.4f %4.4f %d %d %d\n")


def write_cells_text(target, t: float, cells: np.ndarray, stress: np.ndarray) -> None:
    """The output step's cell file (ABM/Simulation.cpp:320-326 with
    ABM/InputOutput.cpp:799-804): one line per cell in ID order, with the
    stress diagonal of mean_stress (stress[:, 0:3], e.g. from
    Colony.mean_stress). `target` is a path (written anew) or an open text
    file (written at its position)."""
    own = not hasattr(target, "write")
    fh = open(target, "w") if own else target
    try:
        for i, c in enumerate(cells):
            p, q, v = c["p"], c["q"], c["Velocity"]
            df, sf = c["DynFric"], c["StaFric"]
            fh.write(CELLS_LINE % (t, i, c["Type"], p[0], p[1], p[2], q[0], q[1], q[2], c["Length"], stress[i, 0],
                                   stress[i, 1], stress[i, 2], v[0], v[1], v[2], c["GrowthRate"], df[0], df[1], df[2],
                                   sf[0], sf[1], sf[2], c["time_p"], c["time_q"], c["age_p"], c["age_q"],
                                   c["Ancestor"]))
        fh.flush()
    finally:
        if own:
            fh.close()


def save_restart(path: str, cells: np.ndarray, t: float, dt: float) -> None:
    """Write a Restart binary as SaveCells does (InputOutput.cpp:434-448)."""
    cells = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
    with open(path, "wb") as fh:
        fh.write(np.array([t, dt], "<f8").tobytes())
        fh.write(cells.tobytes())


RAND_MAX = 2147483647  # glibc


def first_cells(p: Params, colony_size: int = 1, rand=None) -> np.ndarray:
    """The initial colony without a restart, AddFirstCells
    (InputOutput.cpp:294-390): one rod of length L_divide/2 - R lying along x
    through the origin at height R, Ldiv = L + L_divide/2, Type 1, Ancestor 1,
    at rest. Each placement attempt takes one rand() draw (its random radius
    is not used); `rand` is the draw function of the run's stream.

    The reference places every cell of a colony at the same point, so a
    second cell always overlaps the first and the reference gives up after
    10000 attempts ("Unable to generate initial cells!"); only
    IniColonySize 1 starts a run, and other sizes raise ValueError."""
    if int(colony_size) != 1:
        raise ValueError(f"IniColonySize {colony_size}: the reference cannot place more than one initial cell")
    c = np.zeros(1, CELL_DTYPE)
    L = p.L_divide * 0.5 - p.cellRadius
    c["Length"] = L
    c["Radius"] = p.cellRadius
    if rand is not None:
        rand()
    c["p"] = (L / 2.0, 0.0, p.cellRadius)
    c["q"] = (-L / 2.0, 0.0, p.cellRadius)
    c["Ancestor"] = 1
    c["Ldiv"] = L + p.L_divide / 2
    c["Type"] = 1
    return c

