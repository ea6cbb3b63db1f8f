"""Shared test inputs: planar states from the reference's demo restart files
and the demo input parameters, with synthetic injected fields."""
from __future__ import annotations

import os

import numpy as np

from baccolonygrowth_amd.cells import CELL_DTYPE, Params, read_parameters
from baccolonygrowth_amd.colony import synthetic_fields

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DEMO_INPUT = os.path.join(GOLDEN, "inputFile_20mM_SmallSimulation.txt")


def demo_params(max_neighbours: int = 96) -> Params:
    """The demo input (run_sim/inputFile_20mM_SmallSimulation.txt) as the
    reference reads it; BoxLength = L_divide + 2R = 4 um."""
    p, _ = read_parameters(DEMO_INPUT)
    p.maxNeighbours = max_neighbours
    p.planar = 1
    return p


def restart_cells(k: int):
    z = np.load(os.path.join(GOLDEN, f"restart_{k}.npz"))
    cells = np.ascontiguousarray(z["cells"]).view(CELL_DTYPE).reshape(-1).copy()
    return float(z["t"]), float(z["dt"]), cells


def planar_case(k: int, seed: int, divide_frac: float = 0.05):
    """Restart state k with fields and a fraction of cells set to divide
    within the first steps (Ldiv just above the current length)."""
    p = demo_params()
    t, dt, cells = restart_cells(k)
    rng = np.random.default_rng(seed)
    sel = rng.random(len(cells)) < divide_frac
    cells["Ldiv"][sel] = cells["Length"][sel] * (1.0 + rng.uniform(0.0, 5e-5, sel.sum()))
    fields = synthetic_fields(p, cells, seed=seed)
    return p, t, dt, cells, fields


def sorted_lists(counts, lists):
    """Neighbour lists as sorted tuples (set comparison independent of order)."""
    return [tuple(sorted(lists[i, : counts[i]].tolist())) for i in range(len(counts))]


# the restart run of the driver test: the demo input on the reduced grid,
# update and output every 0.001 h, restart index 5
DRIVER_OVERRIDES = {"Box_x": "128", "Box_z": "16", "Box_z_agar": "64", "MaxCells": "4000", "Update_Time": "0.001",
                    "Output_Time": "0.001", "Restart": "1", "RestartIndex": "5"}


def driver_input(path: str, restart_dir: str, t_max: float) -> None:
    """The demo input with the driver run's grid, times and restart lines."""
    over = dict(DRIVER_OVERRIDES, Restart_dir=restart_dir, t_max=repr(t_max))
    out = []
    with open(DEMO_INPUT) as fh:
        for line in fh:
            parts = line.split()
            if parts and parts[0] in over:
                line = f"{parts[0]}\t{over.pop(parts[0])}\n"
            out.append(line)
    out += [f"{k}\t{v}\n" for k, v in over.items()]
    with open(path, "w") as fh:
        fh.writelines(out)
