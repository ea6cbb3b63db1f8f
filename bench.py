"""Benchmark of the mechanics hot path on one MI355X.

Workload: the 2^20-cell kernel bench - spherocylinders (R = 0.5 um, L ~
U[2,4] um, shortened where needed to keep overlaps <= 0.05 um) in a dome of
radius 600 um and height 20 um, tilt ~N(0,15 deg), 3-D model, 5 um bins.
One step = counting-sort bins + full neighbour-list rebuild + both force
passes, integration and growth + division (compaction, divide kernel),
i.e. one cell-update of every cell. Inputs are resident in HBM before the
timed region; the steps are timed between device synchronisations.

Prints one JSON line (rank 0, one GPU):
  value        cell-updates/s over the timed steps
  roofline     the dominant kernel's algorithmic HBM bytes / its mean
               duration (HIP events on the library's stream)
  cpu_baseline the oracle restatement (serial C, FP64, same algorithm) on
               the same colony for a few steps, timed on this host
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s at 1M cells; MG solve ms mixed-vs-FP64 at 1e-10 resid"
HBM_PEAK_GBS = 8000.0
DT_H = 1.0 / 3600.0  # 1 s mechanics step, in the reference's time unit (hours)
GR_BASE = 0.875  # GrowthRate1 such that 1.585 * GR = ln2 / 0.5 h (30 min doubling)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic HBM bytes per cell of each phase (each datum read or written
# once; neighbour records that the gathers re-read are cache traffic, not
# algorithmic). kbar = mean reduced list length.
def phase_bytes(kbar: float) -> dict:
    return {
        # count: p,q (48) + key, slot writes (8) + count RMW (8) + FP32 end
        # points (32); place: key, slot, start (12) + write (4); rank: position,
        # key, bin bounds (16) + sorted (4) + end points in/out (64); scan (~5)
        "bins": 201.0,
        # own p,q,L,R (64) + own id (4) + candidate stream read once (32 + 4)
        # + order/slot/count writes (12) + list writes (4 kbar)
        "lists": 116.0 + 4.0 * kbar,
        # order+count (8) + list (4 kbar) + hot record in/out (128+128) +
        # type, wall, growth rate (20) + DynFric/StaFric (48) + their cell ID (4)
        "move_grow": 336.0 + 4.0 * kbar,
        # the division bitmap, read by the count and the write pass
        "compact": 0.25,
        "divide": 0.0,
    }


# kernel of each timed phase (the roofline's dominant kernel)
PHASE_KERNEL = {"bins": "k_bin_count", "lists": "k_gather_all", "move_grow": "k_move_grow",
                "compact": "k_flag_write", "divide": "k_divide"}
PROFILE = os.path.join(ROOT, "profiles", "bench_default_kernels.txt")


def device_sync() -> None:
    """Wait for all work queued on the device (every stream)."""
    import ctypes

    if ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def device_used_bytes() -> int:
    """Device memory in use (total - free, hipMemGetInfo)."""
    import ctypes

    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if ctypes.CDLL("libamdhip64.so").hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return total.value - free.value


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy calibration) from the rocprofv3 summary of this workload in
    profiles/, or None when the summary has no row for it."""
    try:
        with open(PROFILE) as fh:
            for line in fh:
                f = line.split()
                if len(f) >= 7 and f[0] == kernel:
                    v = (float(f[4]) + float(f[5])) * 1e6
                    return v if v == v else None
    except OSError:
        return None
    return None


def make_workload(n: int, seed: int = 2, radius: float = 600.0):
    from baccolonygrowth_amd.colony import bench_dome

    return bench_dome(n, seed=seed, radius=radius, gr_base=GR_BASE, dt_h=DT_H)


def host_threads() -> int:
    """Threads the CPU baselines use: OMP_NUM_THREADS when set (the host's
    share of cores), else every core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, cells, fields, steps: int, steps_1t: int = 2) -> dict:
    """The oracle (C restatement of the same algorithm, FP64; its OpenMP form:
    list rebuild and move/grow on all threads, bins, divisions and draws in
    order, bit-identical to the serial form) on the same colony: `steps`
    full steps (list rebuild + move/grow + divisions) on every host thread,
    and `steps_1t` on one thread."""
    from oracle.oracle import Oracle

    o = Oracle()
    threads = host_threads()

    def run(nthreads, nsteps):
        o.set_threads(nthreads)
        s = o.sim(p)
        s.set_cells(cells)
        s.set_fields(*fields)
        o.srand(2)
        updates, t = 0, 0.0
        t0 = time.perf_counter()
        for _ in range(nsteps):
            updates += s.n
            s.step_par(DT_H, t, rebuild=True)
            t += DT_H
        el = time.perf_counter() - t0
        s.close()
        return updates / el, el

    v_all, el_all = run(threads, steps)
    log(f"[bench] CPU baseline ({threads} threads): {v_all:.3e} cell-updates/s ({el_all:.1f} s)")
    v_one, el_one = run(1, steps_1t)
    log(f"[bench] CPU baseline (1 thread): {v_one:.3e} cell-updates/s ({el_one:.1f} s)")
    return {
        "value": v_all,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"the same {len(cells)}-cell dome, {steps} full mechanics steps incl. list rebuild (without the "
        f"nutrient V-cycle), the oracle's OpenMP step on {threads} threads ({el_all:.1f} s)",
        "one_thread": {"value": v_one, "cores": 1, "sample": f"{steps_1t} steps on 1 thread ({el_one:.1f} s)"},
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "faithfulness": "the serial oracle step took 0.89x the time of the reference's own step compiled from "
        "source, on the same planar state (1952 cells, 300 steps, one thread, 8-core Xeon build host)",
    }


# 3-D nutrient step (513 x 513 x 129 vertices at 1 um; D = 600 um^2/s,
# dt = 10 s, K = 0.1 mM, reservoir 20 mM) with the Monod sink of the cells
# of a 2^22-cell dome of radius 1200 um that fall inside the grid.
MG_SHAPE = (129, 513, 513)
MG_H = 1.0
Q_CELL = 1.66  # uptake per cell, mM um^3 / s (about 1e6 glucose molecules/s)
# nutrient grid of the coupled step: 257 x 257 x 65 vertices at 2 um, 6 levels,
# centred on the colony (covers |x|, |y| <= 256 um of the 600 um dome)
STEP_SHAPE = (65, 257, 257)
STEP_H = 2.0
STEP_LEVELS = 6


def vcycle_words(levels, nu1: int, nu2: int) -> float:
    """Algorithmic words moved by one V-cycle, summed over levels: a red-black
    sweep reads u, f, a0 and writes u (4 words/point), the residual 4, the
    restriction 1 + 1/8, the prolongation 1/8 + 2, the coarse clear 1/8."""
    words = 0.0
    for lv in levels[:-1]:
        n = float(lv[0]) * lv[1] * lv[2]
        words += n * (4.0 * (nu1 + nu2) + 4.0 + 1.125 + 2.125 + 0.125)
    n = float(levels[-1][0]) * levels[-1][1] * levels[-1][2]
    return words + n * 4.0 * 60  # coarsest: sweeps in FP64


def mg_bench(repeats: int = 3) -> dict:
    from baccolonygrowth_amd.colony import dome_window_centres
    from baccolonygrowth_amd.nutrients import SMOOTH_FUSED_WIDE, Nutrients3D, default_rd3d_params

    nz, ny, nx = MG_SHAPE
    half = (nx - 1) * MG_H / 2
    cm = dome_window_centres(1200.0, 20.0, half, seed=3)
    n_total = 1 << 22
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=MG_H, D=600.0, dt=10.0, K=0.1, c_res=20.0,
                              origin=(-half, -half, 0.0), nu1=2, nu2=2, inner=2)
    res = {}
    fields = {}
    with Nutrients3D(prm) as s:
        inside = s.deposit(cm, np.full(len(cm), Q_CELL))
        levels = s.levels()
        c0 = s.field()
        for mode in ("fp64", "mixed"):
            times = []
            for _ in range(repeats):
                s.set_field(c0)
                it, rr = s.step(mode, tol=1e-10, max_iter=100)
                times.append(s.last_times()["solve"])
            fields[mode] = s.field()
            res[mode] = {"ms": float(np.median(times)), "iterations": it, "relres": rr}
            log(f"[bench] MG {mode}: {res[mode]['ms']:.2f} ms, {it} iterations, relres {rr:.2e}")
        s.set_field(c0)
        vms = {m: float(np.median([s.vcycle(m) for _ in range(5)])) for m in ("fp64", "mixed")}
    # the FP64 solve with its finest-level sweeps on the wide tile too (an
    # option, BCG_RD3D_SMOOTH_FUSED_WIDE; the same iterates, bitwise)
    prm_w = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=MG_H, D=600.0, dt=10.0, K=0.1, c_res=20.0,
                                origin=(-half, -half, 0.0), nu1=2, nu2=2, inner=2, smoother=SMOOTH_FUSED_WIDE)
    with Nutrients3D(prm_w) as s:
        s.deposit(cm, np.full(len(cm), Q_CELL))
        times = []
        for _ in range(repeats):
            s.set_field(c0)
            s.step("fp64", tol=1e-10, max_iter=100)
            times.append(s.last_times()["solve"])
        fp64_wide = float(np.median(times))
    log(f"[bench] MG fp64, wide FP64 tile: {fp64_wide:.2f} ms")
    words = vcycle_words(levels, 2, 2)
    gbs = {m: words * (8 if m == "fp64" else 4) / (vms[m] * 1e-3) / 1e9 for m in vms}
    err = float(np.abs(fields["mixed"] - fields["fp64"]).max())
    return {
        "grid": [nx, ny, nz],
        "h_um": MG_H,
        "levels": [list(map(int, l)) for l in levels],
        "cells_deposited": int(inside),
        "cells_total": n_total,
        "fp64_ms": res["fp64"]["ms"],
        "mixed_ms": res["mixed"]["ms"],
        "speedup": res["fp64"]["ms"] / res["mixed"]["ms"],
        "fp64_wide_tile_ms": fp64_wide,
        "speedup_vs_fp64_wide_tile": fp64_wide / res["mixed"]["ms"],
        "fp64_iterations": res["fp64"]["iterations"],
        "mixed_refinements": res["mixed"]["iterations"],
        "relres": {m: res[m]["relres"] for m in res},
        "mixed_vs_fp64_max_abs_mM": err,
        "vcycle_ms": vms,
        "vcycle_roofline": {
            m: {"achieved": gbs[m], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs[m] / HBM_PEAK_GBS}
            for m in gbs
        },
    }


# Planar nutrient update of the reference's demo input (run_sim/
# inputFile_20mM_SmallSimulation.txt: 256 x 50 colony boxes, 256 x 32 agar
# points on 2 nested levels, 8 multigrid levels, Update_Time 0.01 h -> 266
# outer iterations of 3 V-cycles + 200 colony/wall sweeps) on the demo's
# restart state 8 (1952 cells).
ENV_INPUT = os.path.join(ROOT, "tests", "golden", "inputFile_20mM_SmallSimulation.txt")
ENV_STATE = os.path.join(ROOT, "tests", "golden", "restart_8.npz")


SMALL_SHAPE = (17, 65, 65)
SMALL_H = 2.0


def small_run(until: int = 5000, max_steps: int = 40000, update_every: int = 10) -> dict:
    """The founding-colony run, FP64: 16 rods in a 5 um disk on the agar,
    1 s mechanics steps (lists rebuilt every step), a nutrient update every
    `update_every` steps on 65 x 65 x 17 at 2 um with 4 multigrid levels
    (deposition from the cells, the implicit step solved to 1e-10 in FP64,
    growth rates), until `until` cells."""
    from baccolonygrowth_amd.colony import growing_params, small_colony
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params
    from baccolonygrowth_amd.simulation import Colony

    cells = small_colony(16, seed=1)
    p = growing_params(128.0, 24.0, until + 4096)
    nz, ny, nx = SMALL_SHAPE
    half = (nx - 1) * SMALL_H / 2
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=SMALL_H, D=600.0, dt=update_every * DT_H * 3600.0, K=0.1,
                              c_res=20.0, origin=(-half, -half, 0.0), max_levels=4)
    rate = Q_CELL / 3.0
    with Colony(p) as col, Nutrients3D(prm) as s:
        col.LoadCells(cells)
        col.srand(1)
        s.growth_to_cells(col, GR_BASE, 0.1)
        t, k, updates, cell_updates, its = 0.0, 0, 0, 0, []
        device_sync()
        ts = time.perf_counter()
        while col.n < until and k < max_steps:
            cell_updates += col.n
            col.step(DT_H, t, update=True, no_patch=True)
            t += DT_H
            k += 1
            if k % update_every == 0:
                s.deposit_cells(col, rate)
                it, rr = s.step("fp64", tol=1e-10, max_iter=100)
                its.append(it)
                s.growth_to_cells(col, GR_BASE, 0.1)
                updates += 1
            if k % 2000 == 0:
                log(f"[small] step {k}: {col.n} cells")
        device_sync()
        el = time.perf_counter() - ts
        n_end = col.n
        cmin = float(s.field().min())
    log(f"[small] {n_end} cells after {k} steps ({k / 3600:.2f} h, {updates} nutrient updates): {el:.2f} s")
    return {"cells_start": 16, "cells_end": int(n_end), "steps": k, "hours": k / 3600.0, "nutrient_updates": updates,
            "grid": [nx, ny, nz], "fp64_iterations_mean": float(np.mean(its)) if its else 0.0, "seconds": el,
            "cell_updates_per_s": cell_updates / el, "field_min_mM": cmin}


# V-cycles of the finite-difference forward solves: enough to reach the
# FP64 round-off floor (the FP64 multigrid contracts the residual ~15x per
# cycle on these grids)
FD_VCYCLES = 30


def adjoint_run(n: int = 1 << 18) -> dict:
    """Gradients of one implicit nutrient step on 257 x 257 x 65 at 2 um with
    the sink of n cells (a dome of radius 300 um, rates = base x
    lognormal(0, 0.3)): loss sum (c - c*(r))^2, c* = 20 (1 - 0.8 exp(-r/150))
    mM, with respect to every cell's rate, D and K through
    torch.autograd; checked by central differences (h = 1e-6, 1e-4 and 1e-2
    relative, evaluated in difference form) for D, K and the three cells of
    largest gradient, and the mixed-precision gradient against the FP64 one."""
    import torch

    from baccolonygrowth_amd.adjoint import loss_difference, nutrient_step
    from baccolonygrowth_amd.colony import dome_colony
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params

    cells = dome_colony(n, radius=300.0, height=20.0, seed=4)
    rng = np.random.default_rng(4)
    cm = np.ascontiguousarray((cells["p"] + cells["q"]) * 0.5)
    base = Q_CELL / 3.0 * cells["Length"] * rng.lognormal(0.0, 0.3, n)
    nz, ny, nx = STEP_SHAPE
    half = (nx - 1) * STEP_H / 2
    origin = (-half, -half, 0.0)
    D0, K0 = 600.0, 0.1
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=STEP_H, D=D0, dt=10.0, K=K0, c_res=20.0, origin=origin,
                              max_levels=STEP_LEVELS)
    target = torch.from_numpy(radial_target(STEP_SHAPE, STEP_H, origin))
    c0 = torch.full(STEP_SHAPE, 20.0, dtype=torch.float64)
    out = {"cells": n, "grid": [nx, ny, nz]}
    with Nutrients3D(prm) as s:
        inside = s.deposit(cm, base)

        def loss(rates, D, K, mode):
            c = nutrient_step(rates, D, K, c0, s, cm, mode, 1e-12 if mode == "fp64" else 1e-11)
            return ((c - target) ** 2).sum()

        grads = {}
        for mode in ("fp64", "mixed"):
            r = torch.tensor(base, requires_grad=True)
            Dt = torch.tensor(D0, dtype=torch.float64, requires_grad=True)
            Kt = torch.tensor(K0, dtype=torch.float64, requires_grad=True)
            ta = time.perf_counter()
            L = loss(r, Dt, Kt, mode)
            tf = time.perf_counter()
            L.backward()
            tb = time.perf_counter()
            grads[mode] = (r.grad.numpy().copy(), Dt.grad.item(), Kt.grad.item())
            out[f"{mode}_forward_s"] = tf - ta
            out[f"{mode}_backward_s"] = tb - tf
            log(f"[adjoint] {mode}: forward {tf - ta:.2f} s, backward {tb - tf:.2f} s (host copies included)")
        gr, gD, gK = grads["fp64"]
        out["mixed_vs_fp64_rel"] = {
            "rates": float(np.abs(grads["mixed"][0] - gr).max() / np.abs(gr).max()),
            "D": abs(grads["mixed"][1] - gD) / abs(gD),
            "K": abs(grads["mixed"][2] - gK) / abs(gK),
        }
        pick = np.argsort(-np.abs(gr))[:3]

        c0_np = c0.numpy()
        tgt = target.numpy()

        fd = {}
        # central differences in difference form (loss_difference): each
        # step's field difference solved directly, so that h = 1e-6 is not
        # lost below the solves' round-off floor
        for h in (1e-6, 1e-4, 1e-2):
            e = {}
            for name, v0, g in (("D", D0, gD), ("K", K0, gK)):
                Dm, Dp = (D0 * (1 - h), D0 * (1 + h)) if name == "D" else (D0, D0)
                Km, Kp = (K0 * (1 - h), K0 * (1 + h)) if name == "K" else (K0, K0)
                dL = loss_difference(s, [cm], [base], None, c0_np, tgt, Dm, Dp, Km, Kp, max_iter=FD_VCYCLES)
                e[name] = abs(dL / (2 * h * v0) - g) / abs(g)
            for i in pick:
                rm = base.copy()
                rm[i] *= 1 - h
                dr = np.zeros_like(base)
                dr[i] = 2 * h * base[i]
                dL = loss_difference(s, [cm], [rm], [dr], c0_np, tgt, D0, D0, K0, K0, max_iter=FD_VCYCLES)
                e[f"rate[{int(i)}]"] = abs(dL / (2 * h * base[i]) - gr[i]) / abs(gr[i])
            fd[f"h={h:g}"] = e
            log(f"[adjoint] central differences h={h:g}: " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()))
    out.update({"cells_in_grid": int(inside), "gradient": {"D": gD, "K": gK, "rates_max_abs": float(np.abs(gr).max())},
                "fd_rel_error": fd})
    return out


RUN_CELLS = 1 << 22
RUN_RADIUS = 1200.0


def radial_target(shape, h: float, origin) -> np.ndarray:
    """c(r) = 20 mM (1 - 0.8 exp(-r / 150 um)), r the distance from the
    colony axis, on every plane of the grid."""
    nz, ny, nx = shape
    x = origin[0] + h * np.arange(nx)
    y = origin[1] + h * np.arange(ny)
    r = np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)
    return np.ascontiguousarray(np.broadcast_to(20.0 * (1.0 - 0.8 * np.exp(-r / 150.0)), shape))


def coupled_run(steps: int = 200, update_every: int = 10, ckpt: int = 5) -> dict:
    """The full coupled run: 2^22 cells from the dome scaled to r = 1200 um,
    FP64 mechanics (1 s steps, lists rebuilt every step) with a nutrient
    update every `update_every` steps on 513 x 513 x 129 at 1 um (deposition
    from the cells, the implicit 10 s step by the mixed-precision solver to
    1e-10, growth rates from the field). Then the loss ||c - c_target||^2 of
    the last field is backpropagated through all the updates (checkpoints
    every `ckpt`, recomputation in between; forward and adjoint solves to
    1e-11) to D, K and the cells' rates, checked by central differences
    (h = 1e-6 and 1e-4 relative, evaluated in difference form)."""
    import torch

    from baccolonygrowth_amd.adjoint import rollout_seq
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params
    from baccolonygrowth_amd.simulation import Colony

    used0 = device_used_bytes()
    t0 = time.time()
    p, cells, fields = make_workload(RUN_CELLS, radius=RUN_RADIUS)
    log(f"[run] {len(cells)} cells (r = {RUN_RADIUS:.0f} um), built in {time.time() - t0:.1f}s")
    nz, ny, nx = MG_SHAPE
    half = (nx - 1) * MG_H / 2
    origin = (-half, -half, 0.0)
    D0, K0 = 600.0, 0.1
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=MG_H, D=D0, dt=update_every * DT_H * 3600.0, K=K0, c_res=20.0,
                              origin=origin, nu1=2, nu2=2, inner=2)
    rate = Q_CELL / 3.0  # per um of cell length
    lo, hi = np.array(origin), np.array(origin) + MG_H * (np.array([nx, ny, nz]) - 1)
    seq_cm, seq_rate, iters = [], [], []
    with Colony(p) as col, Nutrients3D(prm) as s:
        col.LoadCells(cells)
        col.set_fields(*fields)
        col.srand(2)
        c_init = s.field()
        peak = device_used_bytes()
        t = 0.0
        cell_updates = 0
        excl = 0.0
        device_sync()
        ts = time.perf_counter()
        for k in range(steps):
            cell_updates += col.n
            col.step(DT_H, t, update=True, no_patch=True)
            t += DT_H
            if (k + 1) % update_every:
                continue
            s.deposit_cells(col, rate)
            it, _ = s.step("mixed", tol=1e-10, max_iter=100)
            iters.append(it)
            s.growth_to_cells(col, GR_BASE, K0)
            # this update's sources for the backward pass (not timed)
            device_sync()
            a = time.perf_counter()
            cc = col.cells()
            cm = (cc["p"] + cc["q"]) * 0.5
            g = np.floor((cm - lo) / MG_H)
            inside = np.all((g >= 0) & (g < np.array([nx, ny, nz]) - 1), axis=1)
            seq_cm.append(np.ascontiguousarray(cm[inside]))
            seq_rate.append(rate * cc["Length"][inside])
            peak = max(peak, device_used_bytes())
            del cc
            device_sync()
            excl += time.perf_counter() - a
        device_sync()
        el = time.perf_counter() - ts - excl
        n_end = col.n
        c_run = s.field()
        log(f"[run] {steps} steps, {len(iters)} nutrient updates: {el:.2f} s, {cell_updates / el:.3e} cell-updates/s")

        target = torch.from_numpy(radial_target(MG_SHAPE, MG_H, origin))
        base = [torch.from_numpy(r) for r in seq_rate]
        c0 = torch.from_numpy(c_init)

        def loss_of(D, K, a):
            # 1e-11: the mixed solver's floor on 34M points is ~4e-12
            c = rollout_seq([a * r for r in base], D, K, c0, s, seq_cm, checkpoint_every=ckpt, mode="mixed", tol=1e-11,
                            adj_tol=1e-11)
            return c, ((c - target) ** 2).sum()

        Dt = torch.tensor(D0, dtype=torch.float64, requires_grad=True)
        Kt = torch.tensor(K0, dtype=torch.float64, requires_grad=True)
        at = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
        ta = time.perf_counter()
        c_fin, L = loss_of(Dt, Kt, at)
        L.backward()
        t_adj = time.perf_counter() - ta
        peak = max(peak, device_used_bytes())
        grads = {"D": Dt.grad.item(), "K": Kt.grad.item(), "rate_factor": at.grad.item()}
        log(f"[run] forward + backward through {len(iters)} updates: {t_adj:.1f} s; dL/dD {grads['D']:.6e}, "
            f"dL/dK {grads['K']:.6e}, dL/d(rate factor) {grads['rate_factor']:.6e}")
        # central differences at two relative steps, 1e-6 and 1e-4, of the
        # FP64 chain of the 20 updates in difference form (loss_difference:
        # each update's field difference solved directly, to round-off
        # relative to itself)
        from baccolonygrowth_amd.adjoint import loss_difference

        tgt = target.numpy()
        fd = {1e-6: {}, 1e-4: {}}
        for rel_h in fd:
            for name, v0 in (("D", D0), ("K", K0), ("rate_factor", 1.0)):
                Dm, Dp = (D0 * (1 - rel_h), D0 * (1 + rel_h)) if name == "D" else (D0, D0)
                Km, Kp = (K0 * (1 - rel_h), K0 * (1 + rel_h)) if name == "K" else (K0, K0)
                if name == "rate_factor":
                    rm = [r * (1 - rel_h) for r in seq_rate]
                    dr = [r * (2 * rel_h) for r in seq_rate]
                else:
                    rm, dr = seq_rate, None
                dL = loss_difference(s, seq_cm, rm, dr, c_init, tgt, Dm, Dp, Km, Kp, max_iter=FD_VCYCLES)
                fd[rel_h][name] = dL / (2.0 * rel_h * v0)
                log(f"[run] dL/d{name}: adjoint {grads[name]:.8e}, central difference (h {rel_h:g}) "
                    f"{fd[rel_h][name]:.8e}")
    err = {h: {k: abs(grads[k] - fd[h][k]) / abs(fd[h][k]) for k in grads} for h in fd}
    return {
        "cells": RUN_CELLS,
        "cells_end": int(n_end),
        "radius_um": RUN_RADIUS,
        "mechanics_steps": steps,
        "nutrient_updates": len(iters),
        "grid": [nx, ny, nz],
        "mixed_refinements_per_update": iters,
        "seconds": el,
        "cell_updates_per_s": cell_updates / el,
        "ms_per_mechanics_step_incl_updates": 1000.0 * el / steps,
        "cells_in_grid_last_update": int(len(seq_cm[-1])),
        "run_vs_recomputed_field_max_abs_mM": float(np.abs(c_fin.detach().numpy() - c_run).max()),
        "loss": float(L.detach()),
        "gradient": grads,
        "central_difference": {f"h={h:g}": fd[h] for h in fd},
        "gradient_rel_error": {f"h={h:g}": err[h] for h in err},
        "adjoint_s": t_adj,
        "checkpoint_every": ckpt,
        "device_memory_peak_gb": peak / 1e9,
        "device_memory_before_gb": used0 / 1e9,
    }


def _env_case():
    from baccolonygrowth_amd.cells import CELL_DTYPE, read_parameters
    from baccolonygrowth_amd.environment import env_params_from_file

    p, _ = read_parameters(ENV_INPUT)
    p.maxNeighbours, p.planar, p.maxCells = 96, 1, 5000
    ep = env_params_from_file(ENV_INPUT)
    z = np.load(ENV_STATE)
    cells = np.ascontiguousarray(z["cells"]).view(CELL_DTYPE).reshape(-1).copy()
    return p, ep, cells


def env_bench(repeats: int = 3, cpu: bool = True, exec_flags: int = 0) -> dict:
    """One nutrient update (UPDATE density pass + UpdateEnvArrayMG_OxygenAcetate)
    on the GPU, median of `repeats` from the same starting state, and the
    reference's own code (compiled from source into oracle/_ref) on one host
    thread when that library is present. exec_flags: the Environment's EXEC_*
    flags (e.g. 2, EXEC_NO_GRAPH, every phase launched directly: what a
    kernel trace can follow)."""
    from baccolonygrowth_amd.environment import ENV_INIT, ENV_UPDATE, Environment
    from baccolonygrowth_amd.simulation import Colony

    p, ep, cells = _env_case()
    out = {"grid": {"colony": [ep.BoxX, ep.BoxZ], "agar": [ep.BoxX, ep.BoxZAgar], "nested_levels": ep.maxLevelsMG},
           "cells": int(len(cells)), "outer_iterations": int(3600 * ep.UpdateTime * 740 / 100)}
    with Colony(p) as col:
        col.LoadCells(cells)
        with Environment(col, ep) as env:
            if exec_flags:
                env.set_execution(exec_flags)
            env.fields(ENV_INIT)
            env.nutrients()
            st0 = env.state()
            times = []
            for _ in range(repeats):
                env.set_state(**st0)
                t0 = time.perf_counter()
                env.fields(ENV_UPDATE)
                sweeps = env.nutrients()
                times.append(time.perf_counter() - t0)
    out["gpu_ms"] = 1000.0 * float(np.median(times))
    out["sweeps"] = int(sweeps)
    log(f"[bench] planar nutrient update: GPU {out['gpu_ms']:.1f} ms")
    out["cpu_baseline"] = None
    if cpu:
        try:
            from oracle.oracle import RefEnv, RefLib, ref_available
        except ImportError:
            ref_available = None
        if ref_available is not None and ref_available():
            ref = RefLib()

            def ref_update(threads):
                ref.set_threads(threads)
                sim = ref.sim(p)
                sim.set_cells(cells)
                renv = RefEnv(ref, sim, ep)
                renv.fields(ENV_INIT)
                renv.nutrients()
                t0 = time.perf_counter()
                renv.fields(ENV_UPDATE)
                renv.nutrients()
                el = time.perf_counter() - t0
                renv.close()
                sim.close()
                return el

            el1 = ref_update(1)
            threads = host_threads()
            eln = ref_update(threads)
            ref.set_threads(1)
            out["cpu_baseline"] = {"ms": 1000.0 * eln, "cores": threads, "kind": "reference",
                                   "sample": "one update of the same state, the reference's sources compiled with "
                                   "g++ -O3 -fopenmp, its OpenMP sweeps on every host thread (their results race; "
                                   "timing only)",
                                   "one_thread_ms": 1000.0 * el1, "cpu_model": cpu_model()}
            out["speedup_vs_cpu"] = 1000.0 * eln / out["gpu_ms"]
            out["speedup_vs_cpu_one_thread"] = 1000.0 * el1 / out["gpu_ms"]
            log(f"[bench] planar nutrient update: reference CPU {1000.0 * el1:.0f} ms (1 thread), "
                f"{1000.0 * eln:.0f} ms ({threads} threads)")
    return out


LARGE_INPUT = os.path.join(ROOT, "tests", "golden", "inputFile_20mM_LargeSimulation.txt")


def env_bench_large(cpu: bool = True, cpu_outer: int = 2) -> dict:
    """The reference's second shipped input (1520 x 100 colony boxes, Box_z_agar
    512, 3 nested agar levels, Update_Time 0.01 h: 266 outer iterations of 3
    V-cycles + 200 sweeps) on the demo restart state 8: one full nutrient
    update on the GPU, and the reference's own code on `cpu_outer` outer
    iterations (the per-iteration work is the same in every iteration) on one
    host thread and on all of them. Reported per outer iteration and per
    update."""
    from baccolonygrowth_amd.cells import CELL_DTYPE, read_parameters
    from baccolonygrowth_amd.environment import ENV_INIT, ENV_UPDATE, Environment, env_params_from_file
    from baccolonygrowth_amd.simulation import Colony

    p, _ = read_parameters(LARGE_INPUT)
    p.maxNeighbours, p.planar, p.maxCells = 96, 1, 4000
    ep = env_params_from_file(LARGE_INPUT)
    z = np.load(ENV_STATE)
    cells = np.ascontiguousarray(z["cells"]).view(CELL_DTYPE).reshape(-1).copy()
    n_outer = int(3600 * ep.UpdateTime * 740 / 100)
    out = {"grid": {"colony": [ep.BoxX, ep.BoxZ], "agar": [ep.BoxX, ep.BoxZAgar], "nested_levels": ep.maxLevelsMG},
           "cells": int(len(cells)), "outer_iterations": n_outer}
    with Colony(p) as col:
        col.LoadCells(cells)
        with Environment(col, ep) as env:
            env.fields(ENV_INIT)
            env.nutrients()
            env.fields(ENV_UPDATE)
            device_sync()
            t0 = time.perf_counter()
            env.nutrients()
            el = time.perf_counter() - t0
    out["gpu_ms"] = 1000.0 * el
    out["gpu_ms_per_outer_iteration"] = 1000.0 * el / n_outer
    log(f"[bench] LargeSimulation nutrient update: GPU {1000.0 * el:.0f} ms ({n_outer} outer iterations)")
    out["cpu_baseline"] = None
    if cpu:
        try:
            from oracle.oracle import RefEnv, RefLib, ref_available
        except ImportError:
            ref_available = None
        if ref_available is not None and ref_available():
            ref = RefLib()
            ep_cpu = env_params_from_file(LARGE_INPUT)
            ep_cpu.UpdateTime = (cpu_outer + 0.5) * 100 / (3600 * 740)

            def ref_update(threads):
                ref.set_threads(threads)
                sim = ref.sim(p)
                sim.set_cells(cells)
                renv = RefEnv(ref, sim, ep_cpu)
                renv.fields(ENV_INIT)
                renv.nutrients()
                renv.fields(ENV_UPDATE)
                t0 = time.perf_counter()
                renv.nutrients()
                el = time.perf_counter() - t0
                renv.close()
                sim.close()
                return el / cpu_outer

            one = ref_update(1)
            threads = host_threads()
            many = ref_update(threads)
            ref.set_threads(1)
            out["cpu_baseline"] = {
                "ms_per_outer_iteration": 1000.0 * many, "cores": threads, "kind": "reference",
                "one_thread_ms_per_outer_iteration": 1000.0 * one,
                "sample": f"{cpu_outer} outer iterations of the same update, the reference's sources compiled with "
                "g++ -O3 -fopenmp (its threaded sweeps race; timing only)", "cpu_model": cpu_model()}
            out["speedup_vs_cpu"] = many / (el / n_outer)
            out["speedup_vs_cpu_one_thread"] = one / (el / n_outer)
            log(f"[bench] LargeSimulation reference CPU per outer iteration: {1000.0 * one:.0f} ms (1 thread), "
                f"{1000.0 * many:.0f} ms ({threads} threads); GPU {1000.0 * el / n_outer:.1f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mg", action="store_true")
    ap.add_argument("--no-env", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the founding-colony, adjoint and full coupled configurations")
    ap.add_argument("--env-exec", type=int, default=0,
                    help="EXEC_* flags of the planar nutrient update (--only env), e.g. 2: no graph replay")
    ap.add_argument("--only", choices=("mg", "env", "env_large", "coupled", "small", "adjoint"), default=None,
                    help="run one secondary section alone and print its JSON object (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is not built.", flush=True)
        sys.exit(2)

    if args.only == "env":
        print(json.dumps(env_bench(cpu=not args.no_cpu, exec_flags=args.env_exec)), flush=True)
        return
    if args.only == "env_large":
        print(json.dumps(env_bench_large(cpu=not args.no_cpu)), flush=True)
        return
    if args.only == "mg":
        print(json.dumps(mg_bench()), flush=True)
        return
    if args.only == "adjoint":
        print(json.dumps({"adjoint_run": adjoint_run()}), flush=True)
        return
    if args.only == "small":
        print(json.dumps({"small_run": small_run()}), flush=True)
        return
    if args.only == "coupled":
        print(json.dumps({"coupled_run": coupled_run()}), flush=True)
        return

    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params
    from baccolonygrowth_amd.simulation import Colony

    t0 = time.time()
    p, cells, fields = make_workload(args.cells)
    log(f"[bench] workload: {len(cells)} cells, bins {p.BoxX}x{p.BoxY}x{p.BoxZ}, built in {time.time() - t0:.1f}s")
    col = Colony(p)
    col.LoadCells(cells)
    col.set_fields(*fields)
    col.srand(2)
    # the coupled step: mechanics (list rebuild every step), then the Monod
    # sink deposited from the cells, one FP64 V-cycle of the implicit
    # nutrient step (dt = the mechanics step) and the cells' growth rates
    # from the new field, all on the device
    nz, ny, nx = STEP_SHAPE
    half = (nx - 1) * STEP_H / 2
    nut = Nutrients3D(default_rd3d_params(nx=nx, ny=ny, nz=nz, h=STEP_H, D=600.0, dt=DT_H * 3600.0, K=0.1,
                                          c_res=20.0, origin=(-half, -half, 0.0), max_levels=STEP_LEVELS))
    rate = Q_CELL / 3.0  # per um of cell length (mean length 3 um)
    inside = nut.deposit_cells(col, rate, count=True)

    def coupled_step(t):
        # the nutrient update is queued behind the mechanics and the next
        # mechanics step behind it (no host synchronisation between them)
        nd = len(col.step(DT_H, t, update=True, no_patch=True))
        nut.update_cells(col, rate, GR_BASE, 0.1, 1)
        return nd

    t = 0.0
    for _ in range(args.warmup):
        coupled_step(t)
        t += DT_H
    log(f"[bench] warmup done ({args.warmup} steps)")
    device_sync()
    phases = {k: 0.0 for k in ("bins", "lists", "move_grow", "compact", "divide")}
    cell_updates = 0
    maxlist = 0
    ndiv = 0
    ts = time.perf_counter()
    for _ in range(args.steps):
        cell_updates += col.n
        ndiv += coupled_step(t)
        t += DT_H
        for k, v in col.last_step_times().items():
            phases[k] += v
        maxlist = max(maxlist, col.last_step_stats()["max_list"])
    device_sync()  # the last nutrient update is still queued
    el = time.perf_counter() - ts
    vcycle_ms = float(np.median([nut.vcycle("fp64") for _ in range(5)]))  # reported, not part of the timed steps
    counts, _ = col.neighbours(stride=1)
    kbar = float(counts.mean())
    n_end = col.n
    nut.close()
    col.close()
    value = cell_updates / el
    ms_step = 1000.0 * el / args.steps
    avg = {k: v / args.steps for k, v in phases.items()}
    dom = max(avg, key=avg.get)
    nmean = cell_updates / args.steps
    bpc = phase_bytes(kbar)
    achieved = bpc[dom] * nmean / (avg[dom] * 1e-3) / 1e9
    log(f"[bench] {args.steps} steps: {ms_step:.3f} ms/step, phases(ms) " + ", ".join(f"{k}={v:.3f}" for k, v in avg.items())
        + f"; one FP64 V-cycle of the nutrient step alone {vcycle_ms:.3f}")
    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: {args.cpu_steps} oracle steps ...")
        cpu = cpu_baseline(p, cells, fields, args.cpu_steps)
        log(f"[bench] CPU baseline {cpu['value']:.3e} cell-updates/s")
    mg = None
    if not args.no_mg:
        log("[bench] multigrid 513x513x129 ...")
        mg = mg_bench()
    envb = envl = None
    if not args.no_env:
        log("[bench] planar nutrient update (demo grid) ...")
        envb = env_bench(cpu=not args.no_cpu)
        log("[bench] planar nutrient update (LargeSimulation grid) ...")
        envl = env_bench_large(cpu=not args.no_cpu)
    extra = {}
    if not args.no_configs:
        # the other configurations of the benchmark, each measured in full; a
        # failure is recorded, not raised, so that the line above still prints
        for key, fn in (("small_run", small_run), ("adjoint_run", adjoint_run), ("coupled_run", coupled_run)):
            log(f"[bench] {key} ...")
            t0 = time.time()
            try:
                extra[key] = fn()
            except Exception as e:  # noqa: BLE001
                extra[key] = {"error": f"{type(e).__name__}: {e}"}
                log(f"[bench] {key} failed: {e}")
            log(f"[bench] {key} done in {time.time() - t0:.1f} s")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "coupled step: mechanics of a 2^20-cell 3-D dome (r=600 um, h=20 um, 5 um bins, list "
            "rebuild every step) + Monod sink deposited from the cells + one FP64 V-cycle of the implicit nutrient "
            "step on 257x257x65 vertices at 2 um (6 levels) + growth rates from the field",
            "nutrient_grid": [nx, ny, nz],
            "cells_in_nutrient_grid": int(inside),
            "nutrient_vcycle_ms": vcycle_ms,
            "cells": int(args.cells),
            "cells_end": int(n_end),
            "divisions": int(ndiv),
            "bins": [p.BoxX, p.BoxY, p.BoxZ],
            "mean_list_length": kbar,
            "max_list_length": int(maxlist),
            "phase_ms": avg,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(PHASE_KERNEL[dom]),
            "traffic_source": os.path.relpath(PROFILE, ROOT),
            "bytes_per_cell": bpc[dom],
        },
        "cpu_baseline": cpu,
        "mg_solve": mg,
        "nutrient_update": envb,
        "nutrient_update_large": envl,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
