"""The reference's executable as one call: read an input file, start from
one cell or from a restart, run the time loop on the GPU and write the
reference's output tree.

    python -m baccolonygrowth_amd.driver <input file> <threads> <output dir> [--seed S] [--t-max T]

mirrors `CellMD3D <input file> <threads> <output dir>` (CellMD3D.cpp:21-170)
and its RunSimulation (Simulation.cpp:42-433); <threads> is accepted and
ignored (the work runs on the GPU). The reference seeds rand() from the
clock; --seed fixes the stream instead.

Order of a run, as the reference's:
  1. parameters (ReadParameters); rand() seeded;
  2. cells: LoadCells of <Restart_dir>/Restart/<RestartIndex>.txt, with the
     fields of the same index (load_field_restart), or AddFirstCells (one
     draw);
  3. the rough wall, one draw per column (Simulation.cpp:160-168);
  4. output step 0 of the initial pass: INIT fields, Density file, nutrient
     update, Environment, AgarField and WallField files
     (Simulation.cpp:184-226);
  5. while t <= t_max: bcg_sim_prepare; at OutFlag the files of output step
     k = 0, 1, ...: Cells from mean_stress of the cells before the move;
     then bcg_sim_advance, whose moved cells are the Restart file; Density,
     Environment, AgarField and WallField; divisions to lineage/0.txt.

One difference: the reference's initial pass also writes Cells/0.txt with
each cell's surface tension force in the stress columns
(Simulation.cpp:228-239); this driver leaves that file empty. The
reference's own first output step of the loop rewrites output step 0, so the
two trees differ only for runs shorter than one Output_Time.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from .cells import RAND_MAX, first_cells, load_restart, read_parameters
from .environment import ENV_INIT, Environment, clock, env_params_from_file, sim_advance, sim_prepare
from .output import LineageFile, OutputFiles, load_field_restart
from .simulation import Colony

# Constants.cpp:18-22 (OutputTime is 100 * the default initial_dt, fixed at
# start-up whatever dt the input sets)
T_MAX_DEFAULT = 20.0
DT_DEFAULT = 0.00001
OUTPUT_TIME_DEFAULT = 100 * 0.00001


def rough_wall(p, rand) -> np.ndarray:
    """Wall roughness of every column, ((float)rand()/RAND_MAX - 0.5) *
    wall_rough (Simulation.cpp:160-168): the draw and RAND_MAX as float, their
    quotient rounded to float, the rest in double."""
    w = np.zeros(p.BoxX * p.BoxY)
    for i in range(w.size):
        q = np.float32(rand()) / np.float32(RAND_MAX)
        w[i] = (float(q) - 0.5) * p.wall_rough
    return w


def _start_fields(col, p, wall):
    """Height at the cell radius and normals (0, 0, -1) where the first
    passes do not write them, and the wall."""
    nh = p.BoxX * p.refinementGridHeight * p.BoxY * p.refinementGridHeight
    normal = np.zeros((nh, 3))
    normal[:, 2] = -1.0
    col.set_fields(height=np.full(nh, p.cellRadius), normal=normal, wall=wall)


def _write_fields(files: OutputFiles, env: Environment, ep, density: bool = True, rest: bool = True):
    st = env.state() if rest else None
    if density:
        files.density(env.get_fields()["density_shift"], ep.BoxX, ep.BoxZ, ep.BoxY)
    if rest:
        files.environment(st["colony"], ep.BoxX, ep.BoxZ, ep.BoxY)
        files.fields(st["agar"], st["wall"], ep.BoxX, ep.BoxZAgar, ep.maxLevelsMG, ep.BoxY)


def run(input_path: str, out_dir: str, seed: int | None = None, t_max: float | None = None,
        max_neighbours: int = 96, log=print) -> dict:
    """One run of the reference's simulation from `input_path` into
    `out_dir`. Returns a summary: steps, output steps, number of cells,
    divisions, the final time and the final cell records."""
    p, extra = read_parameters(input_path)
    p.maxNeighbours = max_neighbours
    ep = env_params_from_file(input_path)
    t_max = float(extra.get("t_max", T_MAX_DEFAULT)) if t_max is None else float(t_max)
    dt = float(extra.get("dt", DT_DEFAULT))
    output_time = float(extra.get("Output_Time", OUTPUT_TIME_DEFAULT))
    restart = int(extra.get("Restart", 0)) != 0
    seed = int(time.time()) & 0xFFFFFFFF if seed is None else int(seed)
    os.makedirs(out_dir, exist_ok=True)
    summary = {"seed": seed, "steps": 0, "outputs": [], "divisions": 0}
    with Colony(p) as col, LineageFile(out_dir) as lineage:
        col.srand(seed)
        t0 = 0.0
        if restart:
            rdir, ridx = str(extra.get("Restart_dir", "")), int(extra.get("RestartIndex", 0))
            t0, dt, cells = load_restart(os.path.join(rdir, "Restart", f"{ridx}.txt"))
        else:
            cells = first_cells(p, int(extra.get("IniColonySize", 16)), rand=col.rand)
        col.LoadCells(cells)
        with Environment(col, ep) as env:
            if restart:
                st = load_field_restart(rdir, ridx, env.state(), ep.BoxX, ep.BoxZ, ep.BoxZAgar, ep.maxLevelsMG,
                                        ep.BoxY)
                env.set_state(**st)
            _start_fields(col, p, rough_wall(p, col.rand))
            # the initial pass writes output step 0
            with OutputFiles(out_dir, 0) as files:
                env.fields(ENV_INIT)
                _write_fields(files, env, ep, density=True, rest=False)
                env.nutrients()
                _write_fields(files, env, ep, density=False, rest=True)
            c = clock(t0, dt, output_time, ep.UpdateTime)
            lineage.rewind()
            k_out = 0
            while c.t <= t_max:
                sim_prepare(col, env, c)
                out = bool(c.OutFlag)
                t_step = c.t
                files = None
                if out:
                    files = OutputFiles(out_dir, k_out)
                    k_out += 1
                    stress, _ = col.mean_stress()
                    files.cells(t_step, col.cells(), stress)
                pairs, moved = sim_advance(col, c, moved=out, lineage_cap=max(4096, 2 * p.maxCells))
                if out:
                    files.restart(moved, t_step, c.dt)
                    _write_fields(files, env, ep)
                    files.close()
                    summary["outputs"].append(files.id)
                    if log:
                        log(f"t = {t_step:6f}, {len(moved):5d} cells")
                lineage.write(pairs)
                summary["divisions"] += len(pairs)
                summary["steps"] += 1
            summary["cells"] = col.n
            summary["t"] = c.t
            summary["final_cells"] = col.cells()
    return summary


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("input")
    ap.add_argument("threads", type=int, help="accepted for the reference's command line; not used")
    ap.add_argument("out_dir")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--t-max", type=float, default=None)
    ap.add_argument("--max-neighbours", type=int, default=96)
    a = ap.parse_args(argv)
    print(f"Output saved to directory: {a.out_dir}", flush=True)
    s = run(a.input, a.out_dir, seed=a.seed, t_max=a.t_max, max_neighbours=a.max_neighbours,
            log=lambda m: print(m, flush=True))
    print(f"Done: {s['steps']} steps, {s['cells']} cells, {s['divisions']} divisions, seed {s['seed']}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
