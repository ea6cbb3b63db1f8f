"""The C ABI from compiled C++ (tests/cabi/caller.cpp, built by build() into
tests/cabi/bcg_caller and linked against the in-tree libbcg.so).

CPU: the struct sizes and field offsets the C++ compiler sees in
include/bcg.h equal the Python mirror's (ctypes structures and numpy
records), and the default parameters agree. GPU: the time loop split at its
output point, run by the C++ caller alone, against the reference's own loop
(tests/golden/golden_output_loop.npz): stress at output steps and moved
cells at 1e-10, lineage exactly, final cells at 1e-10, the rand() position
exactly.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from baccolonygrowth_amd._lib import Clock, EnvParams
from baccolonygrowth_amd.cells import CELL_DTYPE, Params
from baccolonygrowth_amd.environment import AGA_DTYPE, ENV_DTYPE
from helpers import GOLDEN, restart_cells

CALLER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cabi", "bcg_caller")
needs_caller = pytest.mark.skipif(not os.path.exists(CALLER), reason="tests/cabi/bcg_caller not built (build())")


def _abi():
    out = subprocess.run([CALLER, "abi"], check=True, capture_output=True, text=True, timeout=60).stdout
    sizes, offs, lines = {}, {}, out.splitlines()
    for ln in lines:
        w = ln.split()
        if w[0] == "sizeof":
            sizes[w[1]] = int(w[2])
        elif w[0] == "offset":
            offs[w[1]] = int(w[2])
    return sizes, offs, [ln for ln in lines if ln.startswith("default")]


@needs_caller
def test_compiled_abi_layout_equals_python_mirror():
    sizes, offs, defaults = _abi()
    assert sizes == {"bcg_cell": CELL_DTYPE.itemsize, "bcg_params": ctypes.sizeof(Params),
                     "bcg_env_params": ctypes.sizeof(EnvParams), "bcg_env_cell": ENV_DTYPE.itemsize,
                     "bcg_aga": AGA_DTYPE.itemsize, "bcg_clock": ctypes.sizeof(Clock)}
    for key, off in offs.items():
        struct, field = key.split(".")
        if struct == "bcg_cell":
            assert CELL_DTYPE.fields[field][1] == off, key
        else:
            cls = {"bcg_params": Params, "bcg_env_params": EnvParams, "bcg_clock": Clock}[struct]
            assert getattr(cls, field).offset == off, key
    from baccolonygrowth_amd.environment import default_env_params
    from baccolonygrowth_amd.simulation import default_params

    p, e = default_params(), default_env_params()
    assert defaults[0] == (f"default cellRadius {p.cellRadius:.17g} L_divide {p.L_divide:.17g} BoxX {p.BoxX} "
                           f"BoxZ {p.BoxZ} maxCells {p.maxCells}")
    assert defaults[1] == (f"default BoxZAgar {e.BoxZAgar} maxLevelsMG {e.maxLevelsMG} UpdateTime "
                           f"{e.UpdateTime:.17g} maxCarbon {e.maxCarbon:.17g}")


@pytest.mark.gpu
@needs_caller
def test_compiled_caller_time_loop_matches_reference(tmp_path):
    from test_env import GOLDEN_ENV, _params

    g = np.load(os.path.join(GOLDEN, "golden_output_loop.npz"))
    update_time = float(g["update_time"])
    p, ep = _params(np.load(GOLDEN_ENV), update_time)
    t0, dt, _ = restart_cells(5)
    cells = np.ascontiguousarray(g["cells_in"]).view(CELL_DTYPE).reshape(-1)
    (tmp_path / "params.bin").write_bytes(ctypes.string_at(ctypes.addressof(p), ctypes.sizeof(p)))
    (tmp_path / "env_params.bin").write_bytes(ctypes.string_at(ctypes.addressof(ep), ctypes.sizeof(ep)))
    (tmp_path / "cells.bin").write_bytes(np.array([t0, dt]).tobytes() + cells.tobytes())
    (tmp_path / "wall.bin").write_bytes(np.ascontiguousarray(g["wall"], np.float64).tobytes())
    (tmp_path / "run.txt").write_text(f"{int(g['seed'])} {int(g['steps'])} {update_time!r} {float(g['output_time'])!r}\n")
    print(f"[gpu] compiled caller: {len(cells)} cells, {int(g['steps'])} steps", flush=True)
    r = subprocess.run([CALLER, "loop", str(tmp_path)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    for k in g["out_steps"]:
        S = np.fromfile(tmp_path / f"out{k}_stress.bin").reshape(-1, 5)
        exp = g[f"out{k}_stress"]
        assert np.abs(S - exp).max() / np.abs(exp).max() <= 1e-10, k
        moved = np.fromfile(tmp_path / f"out{k}_moved.bin", CELL_DTYPE)
        exp = np.ascontiguousarray(g[f"out{k}_moved"]).view(CELL_DTYPE).reshape(-1)
        assert len(moved) == len(exp)
        for f in ("p", "q", "Length"):
            assert np.abs(moved[f] - exp[f]).max() <= 1e-10 * max(np.abs(exp[f]).max(), 1.0), (k, f)
    assert sorted(int(k) for k in g["out_steps"]) == sorted(
        int(f[3:-len("_stress.bin")]) for f in os.listdir(tmp_path) if f.endswith("_stress.bin"))
    lin = np.fromfile(tmp_path / "lineage.bin", np.int32).reshape(-1, 2)
    assert np.array_equal(lin, g["lineage"])
    got = np.fromfile(tmp_path / "cells_final.bin", CELL_DTYPE)
    exp = np.ascontiguousarray(g["cells"]).view(CELL_DTYPE).reshape(-1)
    assert len(got) == len(exp)
    for f in ("p", "q"):
        assert np.abs(got[f] - exp[f]).max() <= 1e-10 * max(np.abs(exp[f]).max(), 1.0), f
    assert np.fromfile(tmp_path / "draws.bin", np.int32).tolist() == g["next_draws"].tolist()
