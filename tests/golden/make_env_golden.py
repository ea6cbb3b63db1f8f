"""Write the golden vectors of the planar environment (density, height,
normal, nutrients) and of a coupled run of the time loop, from the
reference's own code compiled from source into oracle/_ref/libbcgref.so.

Run in a container that has the reference checkout:
    make -C oracle ref && python tests/golden/make_env_golden.py [large|output_loop|driver]
Output (committed): tests/golden/golden_env.npz, golden_env_large.npz (the
reference's second shipped input), golden_output_loop.npz (a time loop
with output steps) and golden_driver.npz (a restart run of the executable's
sequence, with its output files); an argument writes only that one - inputs
(grid, times, seed) and the reference's outputs only. The grid is the demo input's with
BoxX = 128, BoxZ = 16, BoxZAgar = 64 - every multigrid level keeps an exact
2:1 ratio in x and at least two z points, so that the reference reads no
element outside its arrays - and the demo restart state 5 (129 cells).
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from baccolonygrowth_amd.cells import CELL_DTYPE  # noqa: E402
from baccolonygrowth_amd.colony import synthetic_fields  # noqa: E402
from baccolonygrowth_amd.environment import env_params_from_file  # noqa: E402
from helpers import DEMO_INPUT, GOLDEN, demo_params, driver_input, restart_cells  # noqa: E402
from oracle.oracle import RefEnv, RefLib  # noqa: E402

GRID = dict(BoxX=128, BoxZ=16)
AGAR_Z = 64
FIELDS = ("Carbon", "Waste", "Oxygen", "Acetate", "GrowthRate1", "GrowthRate2", "GrCarAer", "GrCarAna", "GrAceAer",
          "qCarmnt", "qAcemnt", "pAcemnt", "qOxymnt", "pH", "Hydrox", "Acetic", "Buffer")
# one nested agar level (maxLevelsMG = 1, the reference's default): BoxZAgar
# >= 2^7 keeps z points on every level
GRID_ML1 = dict(BoxX=128, BoxZ=16, BoxZAgar=128)


def params(update_time: float, grid=None, agar_z=AGAR_Z, levels=None):
    grid = grid or GRID
    p = demo_params()
    p.BoxX, p.BoxZ = grid["BoxX"], grid["BoxZ"]
    p.maxCells = 4000
    ep = env_params_from_file(DEMO_INPUT)
    ep.BoxX, ep.BoxZ, ep.BoxZAgar = grid["BoxX"], grid["BoxZ"], agar_z
    if levels is not None:
        ep.maxLevelsMG = levels
    ep.UpdateTime = update_time
    return p, ep


def pack(prefix: str, out: dict, f: dict, st: dict | None, agar: bool = True, na: int | None = None):
    """Fields, the colony (both buffers), the interface rows and the finest
    agar level (its Waste never changes in this path: not stored)."""
    for k in ("density", "density_shift", "inside", "wall_shift", "height", "normal", "bounds"):
        if f is not None:
            out[f"{prefix}_{k}"] = f[k]
    if st is not None:
        for k in FIELDS:
            out[f"{prefix}_colony_{k}"] = st["colony"][k]
            out[f"{prefix}_colony_old_{k}"] = st["colony_old"][k]
        na = na or GRID["BoxX"] * AGAR_Z
        for k in ("CarbonAgar", "OxygenAgar", "AcetateAgar", "WasteAgar"):
            if agar and k != "WasteAgar":
                out[f"{prefix}_agar_{k}"] = st["agar"][k][:na]
            out[f"{prefix}_wall_{k}"] = st["wall"][k]
            out[f"{prefix}_wall_old_{k}"] = st["wall_old"][k]


def main():
    ref = RefLib()
    out = {"grid": np.array([GRID["BoxX"], GRID["BoxZ"], AGAR_Z])}
    t0, dt, cells = restart_cells(5)

    # 1. passes on a fixed cell state: INIT fields, nutrients, UPDATE fields,
    #    HEIGHT fields, nutrients again (multigrid state carried over)
    p, ep = params(1e-4)  # 2 outer iterations per nutrient update
    sim = ref.sim(p)
    sim.set_cells(cells)
    env = RefEnv(ref, sim, ep)
    env.fields(0)
    pack("init", out, env.get_fields(), None)
    out["n1"] = np.array(env.nutrients())
    pack("nut1", out, None, env.get_state(), agar=False)
    env.fields(1)
    pack("update", out, env.get_fields(), None)
    env.fields(2)
    pack("height", out, env.get_fields(), None)
    out["n2"] = np.array(env.nutrients())
    pack("nut2", out, None, env.get_state())
    env.close()
    sim.close()

    # 2. one nested agar level: INIT fields and one nutrient update
    p, ep = params(1e-4, GRID_ML1, GRID_ML1["BoxZAgar"], levels=1)
    out["ml1_grid"] = np.array([GRID_ML1["BoxX"], GRID_ML1["BoxZ"], GRID_ML1["BoxZAgar"]])
    sim = ref.sim(p)
    sim.set_cells(cells)
    env = RefEnv(ref, sim, ep)
    env.fields(0)
    out["ml1_n"] = np.array(env.nutrients())
    pack("ml1", out, env.get_fields(), env.get_state(), na=GRID_ML1["BoxX"] * GRID_ML1["BoxZAgar"])
    env.close()
    sim.close()

    # 3. the coupled time loop: init pass, then steps with the clock flags
    update_time, steps, seed = 1e-3, 25, 11
    p, ep = params(update_time)
    rng = np.random.default_rng(seed)
    cells = cells.copy()
    sel = rng.random(len(cells)) < 0.1  # these divide within the first steps
    cells["Ldiv"][sel] = cells["Length"][sel] * (1.0 + rng.uniform(0.0, 5e-5, sel.sum()))
    out["loop_cells_in"] = np.ascontiguousarray(cells).view(np.uint8)
    sim = ref.sim(p)
    sim.set_cells(cells)
    _, _, wall, _, _ = synthetic_fields(p, cells, seed=seed)
    sim.set_fields(wall=wall)
    ref.srand(seed)
    env = RefEnv(ref, sim, ep)
    env.fields(0)
    env.nutrients()
    t, upd, nxt_u, hd, nxt_h = t0, True, update_time, True, update_time / 10.0
    lineage = []
    for _ in range(steps):
        rebuild = False
        if upd:
            env.fields(1)
            env.nutrients()
            rebuild = True
        if hd:
            env.fields(2)
        lineage += sim.step(dt, t, rebuild).tolist()
        t += dt
        upd = nxt_u <= 0
        nxt_u = (update_time if upd else nxt_u) - dt
        hd = nxt_h <= 0
        nxt_h = (update_time / 10.0 if hd else nxt_h) - dt
    out["loop_wall"] = wall
    out["loop_seed"] = np.array(seed)
    out["loop_steps"] = np.array(steps)
    out["loop_update_time"] = np.array(update_time)
    out["loop_cells"] = np.ascontiguousarray(sim.cells()).view(np.uint8)
    out["loop_lineage"] = np.array(lineage, dtype=np.int32).reshape(-1, 2)
    out["loop_t"] = np.array(t)
    pack("loop", out, env.get_fields(), env.get_state())
    env.close()
    sim.close()
    path = os.path.join(GOLDEN, "golden_env.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes;", len(lineage), "divisions in the loop")
    large(ref)
    output_loop(ref)


# ---------------------------------------------------------------- large grid
# The reference's second shipped input (run_sim/inputFile_20mM_LargeSimulation
# .txt: 1520 x 100 colony boxes, Box_z_agar 512, 3 nested agar levels) with
# Update_Time 1e-4 h (2 outer iterations per update instead of 266), on the
# demo restart state 8 (1952 cells). Every multigrid level keeps z points
# (512 >= 2^(8 - 3)), so the reference reads nothing outside its arrays.
# The agar levels (1521 x 513 points each) are stored as strided samples and
# seeded weighted sums; the colony as the window around the colony, which
# holds every point the update changes (the rest keeps LocalEnv()).
LARGE_INPUT = os.path.join(GOLDEN, "inputFile_20mM_LargeSimulation.txt")
LARGE_UPDATE_TIME = 1e-4
LARGE_STRIDE = (8, 16)  # agar samples: every 8th z, every 16th x
ALL_FIELDS = FIELDS + ("pH", "Hydrox", "Acetic", "Buffer")


def large_params():
    from baccolonygrowth_amd.cells import read_parameters

    p, _ = read_parameters(LARGE_INPUT)
    p.maxNeighbours, p.planar, p.maxCells = 96, 1, 4000
    ep = env_params_from_file(LARGE_INPUT)
    ep.UpdateTime = LARGE_UPDATE_TIME
    return p, ep


def agar_weights(n: int, k: int) -> np.ndarray:
    """Weights of the k-th checksum of an n-point agar array."""
    return np.random.default_rng(1000 + k).random(n)


def pack_large(prefix: str, out: dict, ep, st: dict, window) -> None:
    x0, x1 = window
    bx, bz, bza, ml = ep.BoxX, ep.BoxZ, ep.BoxZAgar, ep.maxLevelsMG
    wc = agar_weights(bx * bz, 2)
    for k in ALL_FIELDS:
        for b in ("colony", "colony_old"):
            out[f"{prefix}_{b}_{k}"] = st[b][k].reshape(bx, bz)[x0:x1]
            # the whole array (outside the window too) as a weighted sum
            out[f"{prefix}_{b}_{k}_sums"] = np.array([st[b][k] @ wc, np.abs(st[b][k]).sum()])
    for k in ("CarbonAgar", "OxygenAgar", "AcetateAgar", "WasteAgar"):
        for w in ("wall", "wall_old"):
            out[f"{prefix}_{w}_{k}"] = st[w][k]
        if k == "WasteAgar":
            continue
        a = st["agar"][k].reshape(ml, bx, bza)
        out[f"{prefix}_agar_{k}_sample"] = a[:, :: LARGE_STRIDE[1], :: LARGE_STRIDE[0]]
        w0, w1 = agar_weights(bx * bza, 0), agar_weights(bx * bza, 1)
        out[f"{prefix}_agar_{k}_sums"] = np.array([[a[l].ravel() @ w0, a[l].ravel() @ w1, np.abs(a[l]).sum()]
                                                   for l in range(ml)])


def large(ref):
    from helpers import restart_cells as rc

    p, ep = large_params()
    _, _, cells = rc(8)
    out = {"grid": np.array([ep.BoxX, ep.BoxZ, ep.BoxZAgar, ep.maxLevelsMG]),
           "update_time": np.array(LARGE_UPDATE_TIME), "restart": np.array(8)}
    sim = ref.sim(p)
    sim.set_cells(cells)
    env = RefEnv(ref, sim, ep)
    env.fields(0)
    f = env.get_fields()
    for k in ("density_shift", "inside", "wall_shift", "bounds"):
        out[f"init_{k}"] = f[k]
    b = f["bounds"]
    window = (max(int(b[0]) - 8, 0), min(int(b[1]) + 9, ep.BoxX))
    out["window"] = np.array(window)
    out["n1"] = np.array(env.nutrients())
    pack_large("nut1", out, ep, env.get_state(), window)
    env.fields(1)
    env.fields(2)
    out["n2"] = np.array(env.nutrients())
    pack_large("nut2", out, ep, env.get_state(), window)
    env.close()
    sim.close()
    path = os.path.join(GOLDEN, "golden_env_large.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


# ------------------------------------------------------ output steps in the loop
# The time loop with output steps in the reference's order
# (ABM/Simulation.cpp:254-427): UpdateFlag fields, nutrients and the list
# rebuild; HeightDensity fields; at OutFlag mean_stress of every cell (the
# cell file; its F_cw draws come first), then move + grow, the moved cells
# before the divisions (the Restart file), divisions. OutputTime equals
# Update_Time, so output steps fall on update steps (11 and 21).
def output_loop(ref):
    update_time, steps, seed = 1e-3, 25, 13
    p, ep = params(update_time)
    t0, dt, cells = restart_cells(5)
    rng = np.random.default_rng(seed)
    cells = cells.copy()
    sel = rng.random(len(cells)) < 0.1
    cells["Ldiv"][sel] = cells["Length"][sel] * (1.0 + rng.uniform(0.0, 5e-5, sel.sum()))
    out = {"cells_in": np.ascontiguousarray(cells).view(np.uint8), "seed": np.array(seed), "steps": np.array(steps),
           "update_time": np.array(update_time), "output_time": np.array(update_time)}
    sim = ref.sim(p)
    sim.set_cells(cells)
    _, _, wall, _, _ = synthetic_fields(p, cells, seed=seed)
    sim.set_fields(wall=wall)
    ref.srand(seed)
    env = RefEnv(ref, sim, ep)
    env.fields(0)
    env.nutrients()
    t = t0
    upd, nxt_u, hd, nxt_h, outf, nxt_o = True, update_time, True, update_time / 10.0, False, update_time
    lineage, out_steps = [], []
    for k in range(steps):
        if upd:
            env.fields(1)
            env.nutrients()
            sim.rebuild()
        if hd:
            env.fields(2)
        if outf:
            S, F = sim.mean_stress()
            out[f"out{k}_stress"], out[f"out{k}_fnet"] = S, F
        lin, moved = ref.step_split(sim, dt, t, False)
        if outf:
            out[f"out{k}_moved"] = np.ascontiguousarray(moved).view(np.uint8)
            out_steps.append(k)
        lineage += lin.tolist()
        t += dt
        outf = nxt_o <= 0
        nxt_o = (update_time if outf else nxt_o) - dt
        upd = nxt_u <= 0
        nxt_u = (update_time if upd else nxt_u) - dt
        hd = nxt_h <= 0
        nxt_h = (update_time / 10.0 if hd else nxt_h) - dt
    out["wall"] = wall
    out["out_steps"] = np.array(out_steps)
    out["cells"] = np.ascontiguousarray(sim.cells()).view(np.uint8)
    out["lineage"] = np.array(lineage, dtype=np.int32).reshape(-1, 2)
    out["next_draws"] = np.array([ref.rand() for _ in range(4)])
    env.close()
    sim.close()
    path = os.path.join(GOLDEN, "golden_output_loop.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes; output steps", out_steps, ";", len(lineage), "divisions")


def driver_run(ref):
    """A restart run as the reference's executable makes it, from the
    reference's own functions: LoadCells and the field restart from files the
    reference wrote (restart state 5 after one nutrient update), rand()
    seeded, the rough wall, the initial pass, then the loop with its output
    steps (cell files, saved cells) and lineage."""
    import lzma
    import tempfile

    from baccolonygrowth_amd.cells import read_parameters
    from baccolonygrowth_amd.environment import env_params_from_file

    steps, seed = 25, 29
    t0, dt, cells = restart_cells(5)
    rng = np.random.default_rng(seed)
    cells = cells.copy()
    sel = rng.random(len(cells)) < 0.1
    cells["Ldiv"][sel] = cells["Length"][sel] * (1.0 + rng.uniform(0.0, 5e-5, sel.sum()))
    t_max = t0 + (steps - 0.5) * dt
    out = {"seed": np.array(seed), "steps": np.array(steps), "t_max": np.array(t_max)}
    with tempfile.TemporaryDirectory() as tmp:
        rs = os.path.join(tmp, "rs")
        for d in ("Restart", "Environment", "AgarField", "WallField"):
            os.makedirs(os.path.join(rs, d))
        inp = os.path.join(tmp, "input.txt")
        driver_input(inp, rs, t_max)
        p, _ = read_parameters(inp)
        p.maxNeighbours = 96
        ep = env_params_from_file(inp)
        ref.save_cells(os.path.join(rs, "Restart", "5.txt"), cells, t0, dt)
        # the field files of the restart: the reference's state after one update
        sim0 = ref.sim(p)
        sim0.set_cells(restart_cells(5)[2])
        env0 = RefEnv(ref, sim0, ep)
        env0.fields(0)
        env0.nutrients()
        env0.output(os.path.join(tmp, "unused_density.txt"), os.path.join(rs, "Environment", "5.txt"))
        env0.append(os.path.join(rs, "AgarField", "5.txt"), os.path.join(rs, "WallField", "5.txt"))
        env0.close()
        sim0.close()
        for d in ("Restart", "Environment", "AgarField", "WallField"):
            raw = open(os.path.join(rs, d, "5.txt"), "rb").read()
            out[f"rs_{d}"] = np.frombuffer(lzma.compress(raw), np.uint8)
        # the run
        sim = ref.sim(p)
        ref.srand(seed)
        t, dt2, n = ref.load_cells(sim, os.path.join(rs, "Restart", "5.txt"))
        assert (t, dt2, n) == (t0, dt, len(cells))
        env = RefEnv(ref, sim, ep)
        env.field_restart(rs, 5)
        ref.sim_rough_wall(sim)
        env.fields(0)
        env.nutrients()
        update_time = ep.UpdateTime
        upd, nxt_u, hd, nxt_h, outf, nxt_o = True, update_time, True, update_time / 10.0, False, 1e-3
        lineage, out_steps = [], []
        k = 0
        while t <= t_max:
            if upd:
                env.fields(1)
                env.nutrients()
                sim.rebuild()
            if hd:
                env.fields(2)
            if outf:
                path = os.path.join(tmp, f"cells_{k}.txt")
                sim.output_cells(path, t)
                out[f"out{len(out_steps)}_cells_text"] = np.frombuffer(lzma.compress(open(path, "rb").read()), np.uint8)
            lin, moved = ref.step_split(sim, dt, t, False)
            if outf:
                out[f"out{len(out_steps)}_moved"] = np.ascontiguousarray(moved).view(np.uint8)
                out[f"out{len(out_steps)}_t"] = np.array(t)
                out_steps.append(k)
                # the output step's field files (the last step's are kept)
                env.output(os.path.join(tmp, "d.txt"), os.path.join(tmp, "e.txt"))
                env.append(os.path.join(tmp, "a.txt"), os.path.join(tmp, "w.txt"))
                for name, f in (("Density", "d.txt"), ("Environment", "e.txt"), ("AgarField", "a.txt"),
                                ("WallField", "w.txt")):
                    out[f"last_{name}"] = np.frombuffer(lzma.compress(open(os.path.join(tmp, f), "rb").read()),
                                                        np.uint8)
            lineage += lin.tolist()
            t += dt
            k += 1
            outf = nxt_o <= 0
            nxt_o = (1e-3 if outf else nxt_o) - dt
            upd = nxt_u <= 0
            nxt_u = (update_time if upd else nxt_u) - dt
            hd = nxt_h <= 0
            nxt_h = (update_time / 10.0 if hd else nxt_h) - dt
        env.close()
        out["steps_run"] = np.array(k)
        out["out_steps"] = np.array(out_steps)
        out["cells"] = np.ascontiguousarray(sim.cells()).view(np.uint8)
        out["lineage"] = np.array(lineage, dtype=np.int32).reshape(-1, 2)
        out["next_draws"] = np.array([ref.rand() for _ in range(4)])
        sim.close()
    path = os.path.join(GOLDEN, "golden_driver.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes; steps", k, "output steps", out_steps, ";", len(lineage), "divisions")


if __name__ == "__main__":
    if sys.argv[1:] == ["large"]:
        large(RefLib())
    elif sys.argv[1:] == ["output_loop"]:
        output_loop(RefLib())
    elif sys.argv[1:] == ["driver"]:
        driver_run(RefLib())
    else:
        main()
