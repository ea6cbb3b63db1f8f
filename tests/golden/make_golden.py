"""Write the golden vectors of the mechanics path from the reference's own
code, compiled from source into oracle/_ref/libbcgref.so (oracle/Makefile).

Run in a container that has the reference checkout:
    make -C oracle ref && python tests/golden/make_golden.py
Outputs (committed): tests/golden/golden_pairs.npz, golden_step_r7.npz,
golden_step_r8.npz. Each holds inputs and the reference's outputs only.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from baccolonygrowth_amd.cells import CELL_DTYPE  # noqa: E402
from helpers import GOLDEN, demo_params, planar_case  # noqa: E402
from oracle.oracle import RefLib  # noqa: E402


def random_pairs(n: int, seed: int, planar: bool):
    """Cell pairs covering the min_distance branches: random, parallel,
    crossing, end-to-end, overlapping and touching."""
    rng = np.random.default_rng(seed)
    a = np.zeros(n, CELL_DTYPE)
    b = np.zeros(n, CELL_DTYPE)
    for c in (a, b):
        cm = rng.normal(0, 1.2, (n, 3))
        cm[:, 2] += 2.0
        th = rng.uniform(0, 2 * np.pi, n)
        ph = rng.uniform(-0.6, 0.6, n)
        L = rng.uniform(0.5, 4.0, n)
        u = np.stack([np.cos(th) * np.cos(ph), np.sin(th) * np.cos(ph), np.sin(ph)], -1)
        if planar:
            cm[:, 1] = 0
            u = np.stack([np.cos(th), np.zeros(n), np.sin(th)], -1)
        c["p"] = cm + 0.5 * L[:, None] * u
        c["q"] = cm - 0.5 * L[:, None] * u
        c["Length"] = L
        c["Radius"] = rng.choice([0.5, 0.45, 0.55], n)
        c["Velocity"] = rng.normal(0, 1.0, (n, 3))
        c["AngularVelocity"] = rng.normal(0, 1.0, (n, 3))
        if planar:
            c["Velocity"][:, 1] = 0
    # special geometries in the first rows
    k = 0
    for off in (0.0, 0.3, 0.9, 1.0, 1.5):  # parallel, side by side
        a["p"][k], a["q"][k] = (1, 0, 1), (-1, 0, 1)
        b["p"][k], b["q"][k] = (1.5, 0, 1 + off), (-0.5, 0, 1 + off)
        a["Length"][k] = b["Length"][k] = 2.0
        k += 1
    for gap in (-0.2, 0.0, 0.5):  # collinear end to end
        a["p"][k], a["q"][k] = (1, 0, 1), (-1, 0, 1)
        b["p"][k], b["q"][k] = (3 + gap, 0, 1), (1 + gap, 0, 1)
        a["Length"][k] = b["Length"][k] = 2.0
        k += 1
    a["p"][k], a["q"][k] = (1, 0, 1), (-1, 0, 1)  # crossing at centre
    b["p"][k], b["q"][k] = (0, 0, 2), (0, 0, 0.2)
    a["Length"][k], b["Length"][k] = 2.0, 1.8
    return a, b


def main():
    ref = RefLib()
    p = demo_params()
    out = {}
    for tag, planar in (("planar", True), ("3d", False)):
        a, b = random_pairs(400, 11 if planar else 12, planar)
        md = np.stack([ref.min_distance(a[i : i + 1], b[i : i + 1]) for i in range(len(a))])
        fcc = np.stack([ref.F_cc(p, a[i : i + 1], b[i : i + 1]) for i in range(len(a))])
        out[f"{tag}_a"] = a.view(np.uint8).reshape(len(a), 208)
        out[f"{tag}_b"] = b.view(np.uint8).reshape(len(b), 208)
        out[f"{tag}_min_distance"] = md
        out[f"{tag}_F_cc"] = fcc
    # F_cw: cells near the agar, one draw each (the draw value is unused
    # because var_pos = 0 in the demo input)
    a, _ = random_pairs(200, 13, True)
    a["p"][:, 2] = np.random.default_rng(13).uniform(0.2, 0.7, 200)
    a["q"][:, 2] = np.random.default_rng(14).uniform(0.2, 0.7, 200)
    wz = np.random.default_rng(15).uniform(-0.05, 0.05, 200)
    ref.srand(5)
    fcw = np.stack([ref.F_cw(p, a[i : i + 1], wz[i]) for i in range(len(a))])
    out["cw_cells"] = a.view(np.uint8).reshape(len(a), 208)
    out["cw_wall_z"] = wz
    out["cw_out"] = fcw
    np.savez_compressed(os.path.join(GOLDEN, "golden_pairs.npz"), **out)

    for k, steps in ((7, 4), (8, 1)):
        p, t, dt, cells, fields = planar_case(k, seed=k)
        sim = ref.sim(p)
        sim.set_cells(cells)
        sim.set_fields(*[f.reshape(-1) for f in fields])
        sim.rebuild()
        counts0, lists0 = sim.neighbours(stride=p.maxNeighbours)
        F0, T0 = sim.sum_forces()
        seed = 1000 + k
        ref.srand(seed)
        lineage = []
        tt = t
        for s in range(steps):
            lin = sim.step(dt, tt, rebuild=(s == 2))
            lineage.append(np.column_stack([np.full(len(lin), s), lin]))
            tt += dt
        counts1, lists1 = sim.neighbours(stride=p.maxNeighbours)
        after = sim.cells()
        next_draws = np.array([ref.rand() for _ in range(4)])
        np.savez_compressed(
            os.path.join(GOLDEN, f"golden_step_r{k}.npz"),
            cells_in=cells.view(np.uint8).reshape(len(cells), 208),
            t=t,
            dt=dt,
            seed=seed,
            steps=steps,
            rebuild_step=2,
            height=fields[0],
            normal=fields[1],
            wall=fields[2],
            gr1=fields[3],
            gr2=fields[4],
            nl_counts0=counts0,
            nl_lists0=lists0,
            F0=F0,
            T0=T0,
            cells_out=after.view(np.uint8).reshape(len(after), 208),
            lineage=np.concatenate(lineage) if lineage else np.zeros((0, 3), np.int32),
            nl_counts1=counts1,
            nl_lists1=lists1,
            next_draws=next_draws,
        )
        print(f"r{k}: {len(cells)} -> {len(after)} cells, {sum(len(x) for x in lineage)} divisions")
        sim.close()


if __name__ == "__main__":
    main()
