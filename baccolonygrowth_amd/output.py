"""The reference's output tree and its field text formats, written from the
state arrays the library hands back, byte for byte as the reference writes
them; and the readers its field restart uses.

Output tree (CreateOutputFiles, AgentBasedModel-BacColonyGrowth/InputOutput.cpp:40-271):
every output step k makes <dir>/<Name>/k.txt for the 17 names of
OUTPUT_DIRS, all opened "w" except Cells ("a" when appending after a
restart). Only Cells, Restart, Density, Environment, AgarField and WallField
are written (Simulation.cpp:341-367); the other eleven stay empty files. The
lineage file <dir>/lineage/0.txt is opened once for the whole run
(CreateOutputFileLineage, InputOutput.cpp:13-33) and takes one "%d %d\\n"
line, parent then daughter, per division (Simulation.cpp:392).

Text formats ("%6E\\t" per value, "\\n" per row):
  * Density/k.txt   DensityShiftP.Output(f, 1) (Array.h:568-583): z planes
    of y rows of x values, a blank line after each plane;
  * Environment/k.txt   Environment.Output(f, 1) (Array.h:931-1100): one block
    per field of ENV_OUTPUT_FIELDS, z planes of y rows, no blank lines;
  * AgarField/k.txt   FieldAgar[level]->Append(f, 1) for every level
    (Array3D<LocalAga>::Append, Array.h:1447-1494): per level the blocks
    CarbonAgar, AcetateAgar, WasteAgar, OxygenAgar (AGAR_BLOCKS), each z
    planes of y rows, every n-th x and y;
  * WallField/k.txt   FieldWall[level]->Append(f) (Array2D<LocalAga>::Append,
    Array.h:419-457): per level the same four blocks, y rows of x values.

The arrays are the library's layouts, which are the reference's: a 3-D field
is x-major (x * BY * BZ + y * BZ + z), the agar of level l starts at
l * BX * BY * BZAgar, the wall of level l at l * BX * BY.
"""
from __future__ import annotations

import math
import os

import numpy as np

# InputOutput.cpp:40-271, in the order the reference creates them
OUTPUT_DIRS = ("Cells", "Restart", "RoughDensity", "RoughDensity1", "RoughDensity2", "Density", "Density1",
               "Density2", "WallDensity", "WallDensity1", "WallDensity2", "RoughHeight", "Height", "Normal",
               "Environment", "AgarField", "WallField")

# Fields of the environment file, in the order its writer prints them
# (Array3D<LocalEnv>::Output, Array.h:931-1100).
ENV_OUTPUT_FIELDS = ("Carbon", "Acetate", "Waste", "Oxygen", "GrowthRate1", "GrowthRate2", "GrCarAer", "GrCarAna",
                     "GrAceAer", "pH", "Hydrox", "Buffer", "Acetic", "qCarmnt", "qAcemnt", "pAcemnt", "qOxymnt")

# the blocks of an agar or wall append (Array.h:419-457, 1447-1494) and of the
# agar / wall readers (Array.h:460-560, 1286-1390)
AGAR_BLOCKS = ("CarbonAgar", "AcetateAgar", "WasteAgar", "OxygenAgar")
# the four fields Array3D<LocalEnv>::Input reads (Array.h:1176-1285)
ENV_INPUT_FIELDS = ("Carbon", "Acetate", "Waste", "Oxygen")


# ------------------------------------------------------------------ writers
def fmt_e(v: float) -> str:
    """One value as C's printf("%6E") prints it; Python's %-formatting
    agrees (both round the exact binary value correctly) except that glibc
    keeps the sign of a NaN."""
    s = "%6E" % v
    if v != v and math.copysign(1.0, v) < 0:
        s = "%6s" % "-NAN"
    return s


def _row(values) -> str:
    """ "%6E\\t" per value, then "\\n"."""
    vals = np.asarray(values, dtype=np.float64)
    if np.any(np.isnan(vals) & np.signbit(vals)):
        return "".join(fmt_e(v) + "\t" for v in vals.tolist()) + "\n"
    return "".join(["%6E\t" % v for v in vals.tolist()]) + "\n"


def _planes(values, BX: int, BY: int, BZ: int, blank: bool, n: int = 1):
    """Rows of every n-th x value, every n-th y row, z planes, from a flat
    x-major array (x * BY * BZ + y * BZ + z); a blank line after each plane
    when `blank`."""
    v = np.asarray(values, dtype=np.float64).reshape(BX, BY, BZ)
    out = []
    for zi in range(BZ):
        for yi in range(0, BY, n):
            out.append(_row(v[::n, yi, zi]))
        if blank:
            out.append("\n")
    return out


def _sink(target, mode: str):
    """An open file for a path, or the file object itself (left open)."""
    if hasattr(target, "write"):
        return target, False
    return open(target, mode), True


def _write(target, lines, mode: str = "w") -> None:
    fh, own = _sink(target, mode)
    try:
        fh.writelines(lines)
        fh.flush()
    finally:
        if own:
            fh.close()


def write_density_text(target, density_shift: np.ndarray, BX: int, BZ: int, BY: int = 1) -> None:
    """The output step's density file, DensityShiftP.Output(Files.density, 1)
    (Simulation.cpp:198, 355; Array3D<double>::Output, Array.h:568-583).
    `target` is a path or an open text file."""
    _write(target, _planes(density_shift, BX, BY, BZ, True))


def write_env_text(target, colony: np.ndarray, BX: int, BZ: int, BY: int = 1) -> None:
    """The output step's environment file, Environment.Output(Files.env, 1)
    (Simulation.cpp:221, 362): one block per field of ENV_OUTPUT_FIELDS."""
    lines = []
    for f in ENV_OUTPUT_FIELDS:
        lines += _planes(colony[f], BX, BY, BZ, False)
    _write(target, lines)


def agar_append_lines(level: np.ndarray, BX: int, BZA: int, BY: int = 1, n: int = 1):
    """Array3D<LocalAga>::Append(f, n) of one agar level (Array.h:1447-1494):
    the four AGAR_BLOCKS, each z planes of rows of every n-th x and y."""
    lines = []
    for f in AGAR_BLOCKS:
        lines += _planes(level[f], BX, BY, BZA, False, n)
    return lines


def wall_append_lines(level: np.ndarray, BX: int, BY: int = 1):
    """Array2D<LocalAga>::Append(f) of one wall level (Array.h:419-457): the
    four AGAR_BLOCKS, each BY rows of BX values (index x * BY + y)."""
    lines = []
    for f in AGAR_BLOCKS:
        v = np.asarray(level[f], dtype=np.float64).reshape(BX, BY)
        lines += [_row(v[:, yi]) for yi in range(BY)]
    return lines


def write_field_appends(agar_target, wall_target, agar: np.ndarray, wall: np.ndarray, BX: int, BZA: int,
                        levels: int, BY: int = 1) -> None:
    """The output step's AgarField and WallField files: for every level,
    FieldAgar[level]->Append(Files.aga, 1) and FieldWall[level]->Append(Files.wal)
    (Simulation.cpp:222-226, 363-367). `agar` holds levels * BX * BY * BZA
    records, `wall` levels * BX * BY (the library's state layout)."""
    na, nw = BX * BY * BZA, BX * BY
    if agar.size != levels * na or wall.size != levels * nw:
        raise ValueError(f"expected {levels * na} agar and {levels * nw} wall records, got {agar.size}, {wall.size}")
    a, own_a = _sink(agar_target, "w")
    w, own_w = _sink(wall_target, "w")
    try:
        for lv in range(levels):
            a.writelines(agar_append_lines(agar[lv * na:(lv + 1) * na], BX, BZA, BY))
            w.writelines(wall_append_lines(wall[lv * nw:(lv + 1) * nw], BX, BY))
        a.flush()
        w.flush()
    finally:
        if own_a:
            a.close()
        if own_w:
            w.close()


def write_lineage(target, pairs) -> None:
    """One "%d %d\\n" line per division, parent then daughter
    (Simulation.cpp:392), appended to an open file or a path."""
    lines = ["%d %d\n" % (int(a), int(b)) for a, b in np.asarray(pairs, dtype=np.int64).reshape(-1, 2)]
    _write(target, lines, "a")


# ------------------------------------------------------------------ readers
def _values(line: str):
    """The doubles `while (iss >> val)` takes from one line: tokens up to the
    first one that is not a decimal number (libstdc++ reads no nan/inf)."""
    out = []
    for tok in line.split():
        low = tok.lower()
        if "n" in low or "i" in low:  # nan, inf, infinity
            break
        try:
            out.append(float(tok))
        except ValueError:
            break
    return out


def read_blocks(lines, out: np.ndarray, fields, BX: int, BY: int, BZ: int, path: str = "") -> None:
    """Fill `fields` of the x-major records `out` from the next blocks of
    `lines` (an iterator of text lines), as the reference's Input readers do
    (Array.h:460-560 for BZ = 1 walls, 1176-1285 and 1286-1390 for 3-D
    fields): per field, lines fill x = 0, 1, ... of row y of plane z, until
    BY * BZ lines are read. A line with more than BX values, or a block cut
    short by the end of the file, is an error (the reference asserts); an
    empty stream leaves the records as they are."""
    for f in fields:
        col = out[f]
        xi = yi = zi = 0
        for line in lines:
            vals = _values(line)
            if len(vals) > BX:
                raise ValueError(f"{path}: {len(vals)} values in a row of {BX}")
            if vals:
                col[np.arange(len(vals)) * (BY * BZ) + yi * BZ + zi] = vals
            yi += 1
            if yi == BY:
                yi = 0
                zi += 1
            if zi == BZ:
                zi = 0
            if yi == 0 and zi == 0:
                break
        if yi or zi:
            raise ValueError(f"{path}: block {f} ends after part of its rows")


def _lines(path: str):
    """std::getline over a file: lines without their '\\n'."""
    with open(path, "r") as fh:
        for line in fh:
            yield line[:-1] if line.endswith("\n") else line


def load_field_restart(dirname: str, index: int, state: dict, BX: int, BZ: int, BZA: int, levels: int,
                       BY: int = 1) -> dict:
    """The field part of a restart (CellMD3D.cpp:105-125), with the reference's
    own reads: `state` (as Environment.state() returns it) is copied and
    filled from <dirname>/Environment/<index>.txt, AgarField/<index>.txt and
    WallField/<index>.txt.

    As in the reference: Environment takes the file's first four blocks
    (Carbon, Acetate, Waste, Oxygen) and oldEnvironment the next four, which
    the writer filled with GrowthRate1, GrowthRate2, GrCarAer and GrCarAna;
    every level reopens the agar and wall files, so every FieldAgar[i] gets
    level 0 of the file and every oldFieldAgar[i] level 1 (left as it was
    when the file holds one level). Fields not read keep their values."""
    st = {k: v.copy() for k, v in state.items()}
    env = os.path.join(dirname, "Environment", f"{index}.txt")
    it = _lines(env)
    read_blocks(it, st["colony"], ENV_INPUT_FIELDS, BX, BY, BZ, env)
    read_blocks(it, st["colony_old"], ENV_INPUT_FIELDS, BX, BY, BZ, env)
    aga = os.path.join(dirname, "AgarField", f"{index}.txt")
    wal = os.path.join(dirname, "WallField", f"{index}.txt")
    na, nw = BX * BY * BZA, BX * BY
    for lv in range(levels):
        it = _lines(aga)
        it3 = _lines(wal)
        read_blocks(it, st["agar"][lv * na:(lv + 1) * na], AGAR_BLOCKS, BX, BY, BZA, aga)
        read_blocks(it, st["agar_old"][lv * na:(lv + 1) * na], AGAR_BLOCKS, BX, BY, BZA, aga)
        read_blocks(it3, st["wall"][lv * nw:(lv + 1) * nw], AGAR_BLOCKS, BX, BY, 1, wal)
        read_blocks(it3, st["wall_old"][lv * nw:(lv + 1) * nw], AGAR_BLOCKS, BX, BY, 1, wal)
    return st


# ------------------------------------------------------------------ the tree
class OutputFiles:
    """The files of one output step (CreateOutputFiles(k, Files, append),
    InputOutput.cpp:40-271): makes the 17 directories and opens k.txt in
    each; close() is CloseOutputFiles."""

    def __init__(self, dirname: str, output_id: int, append: bool = False):
        self.dir, self.id = dirname, int(output_id)
        self.fh = {}
        for name in OUTPUT_DIRS:
            d = os.path.join(dirname, name)
            os.makedirs(d, exist_ok=True)
            path = os.path.join(d, f"{self.id}.txt")
            mode = "wb" if name == "Restart" else ("a" if (name == "Cells" and append) else "w")
            self.fh[name] = open(path, mode)

    def path(self, name: str) -> str:
        return os.path.join(self.dir, name, f"{self.id}.txt")

    def cells(self, t: float, cells: np.ndarray, stress: np.ndarray) -> None:
        """Output(Files.cells, i, t, cell, stress) of every cell (Simulation.cpp:320-326)."""
        from .cells import write_cells_text

        write_cells_text(self.fh["Cells"], t, cells, stress)

    def restart(self, cells: np.ndarray, t: float, dt: float) -> None:
        """SaveCells(Files.restart, cells, N, t, dt) (InputOutput.cpp:434-448):
        rewind, then t, dt and the cell records."""
        from .cells import CELL_DTYPE

        fh = self.fh["Restart"]
        fh.seek(0)
        fh.write(np.array([t, dt], "<f8").tobytes())
        fh.write(np.ascontiguousarray(cells, dtype=CELL_DTYPE).tobytes())
        fh.flush()

    def density(self, density_shift: np.ndarray, BX: int, BZ: int, BY: int = 1) -> None:
        write_density_text(self.fh["Density"], density_shift, BX, BZ, BY)

    def environment(self, colony: np.ndarray, BX: int, BZ: int, BY: int = 1) -> None:
        write_env_text(self.fh["Environment"], colony, BX, BZ, BY)

    def fields(self, agar: np.ndarray, wall: np.ndarray, BX: int, BZA: int, levels: int, BY: int = 1) -> None:
        write_field_appends(self.fh["AgarField"], self.fh["WallField"], agar, wall, BX, BZA, levels, BY)

    def close(self) -> None:
        for fh in self.fh.values():
            fh.close()
        self.fh = {}

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class LineageFile:
    """<dir>/lineage/<id>.txt, open for the whole run
    (CreateOutputFileLineage / CloseOutputFileLineage, InputOutput.cpp:13-38)."""

    def __init__(self, dirname: str, output_id: int = 0):
        d = os.path.join(dirname, "lineage")
        os.makedirs(d, exist_ok=True)
        self.path = os.path.join(d, f"{int(output_id)}.txt")
        self.fh = open(self.path, "w")

    def rewind(self) -> None:
        """rewind(Files.lineage) before the time loop (Simulation.cpp:253)."""
        self.fh.seek(0)

    def write(self, pairs) -> None:
        write_lineage(self.fh, pairs)

    def close(self) -> None:
        if self.fh:
            self.fh.flush()
            self.fh.close()
            self.fh = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
