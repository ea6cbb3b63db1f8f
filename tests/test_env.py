"""Planar colony environment: density / height / normal fields, the nutrient
update with its nested agar multigrid, the time-loop clock, and a coupled
run of the time loop, against the reference's own code.

The expected values are the reference's outputs (tests/golden/golden_env.npz,
written by tests/golden/make_env_golden.py from oracle/_ref, the reference
compiled from source) on the demo input with BoxX = 128, BoxZ = 16,
BoxZAgar = 64 (and, for one nested agar level, BoxZAgar = 128 with
maxLevelsMG = 1) and the demo restart state 5. Arrays the reference leaves
uninitialised are pinned to 0 on both sides (BoxAverage outside its window,
insideColonyDen one box past the last).

Tolerances: the fields and the nutrient state are FP64 through the same
operations in the same order (no contraction on either side); they are
compared at 1e-13 relative. The coupled run adds the mechanics (positions
at 1e-10, as the mechanics tests)."""
import ctypes
import os
import sys

import numpy as np
import pytest

from baccolonygrowth_amd.cells import CELL_DTYPE
from helpers import DEMO_INPUT, GOLDEN, demo_params, restart_cells

GOLDEN_ENV = os.path.join(GOLDEN, "golden_env.npz")
FIELDS = ("Carbon", "Waste", "Oxygen", "Acetate", "GrowthRate1", "GrowthRate2", "GrCarAer", "GrCarAna", "GrAceAer",
          "qCarmnt", "qAcemnt", "pAcemnt", "qOxymnt")
# the acetate buffer state (acetateBufferSolve): Newton iterates, then
# pH = -log(...)/log(10); device log and the reference's pow(x, 2) (here x*x)
# may differ in the last bit, so compared at 1e-12
BUFFER_FIELDS = ("pH", "Hydrox", "Acetic", "Buffer")
RTOL = 1e-13
RTOL_BUFFER = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(g, update_time, grid="grid", levels=None):
    from baccolonygrowth_amd.environment import env_params_from_file

    bx, bz, bza = (int(v) for v in g[grid])
    p = demo_params()
    p.BoxX, p.BoxZ, p.maxCells = bx, bz, 4000
    ep = env_params_from_file(DEMO_INPUT)
    ep.BoxX, ep.BoxZ, ep.BoxZAgar = bx, bz, bza
    if levels is not None:
        ep.maxLevelsMG = levels
    ep.UpdateTime = update_time
    return p, ep


def _close(got, exp, label, rtol=RTOL):
    got, exp = np.asarray(got, float).reshape(-1), np.asarray(exp, float).reshape(-1)
    assert got.shape == exp.shape, label
    d = np.abs(got - exp)
    scale = max(np.abs(exp).max(), 1e-300)
    bad = ~(d <= rtol * scale)
    if bad.any():
        i = int(np.flatnonzero(bad)[0])
        raise AssertionError(f"{label}: {bad.sum()} of {d.size} differ, max {d.max():.3e} (scale {scale:.3e}); "
                             f"first at {i}: got {got[i]!r} expected {exp[i]!r}")
    return float(d.max() / scale)


# ------------------------------------------------------------------ CPU
def test_golden_env_fixture_is_consistent():
    g = np.load(GOLDEN_ENV)
    bx, bz, bza = (int(v) for v in g["grid"])
    assert g["init_density"].shape == (bx * bz,)
    assert g["nut1_colony_Carbon"].shape == (bx * bz,)
    assert g["nut2_agar_CarbonAgar"].shape == (bx * bza,)
    # the footprint is 1 inside, 0.4 on its halo, 0 elsewhere
    assert set(np.unique(g["init_inside"]).tolist()) <= {0.0, 0.4, 1.0}
    # bounds: occupied boxes padded by 5 in x and z
    assert g["init_bounds"][0] >= 1 and g["init_bounds"][1] <= bx - 2
    # n = int(3600 Update_Time 740 / 100) outer iterations of 200 sweeps
    assert int(g["n1"]) == int(3600 * 1e-4 * 740 / 100) * 200
    # growth rates are the sum of the three pathways where they were updated
    c = {k: g[f"nut1_colony_{k}"] for k in ("GrowthRate1", "GrCarAer", "GrCarAna", "GrAceAer")}
    assert np.array_equal(c["GrowthRate1"], c["GrCarAer"] + c["GrCarAna"] + c["GrAceAer"])
    assert len(g["loop_lineage"]) > 0
    assert g["ml1_agar_CarbonAgar"].shape == (int(g["ml1_grid"][0]) * int(g["ml1_grid"][2]),)
    # the buffer state is a pH where it was computed and the LocalEnv() default 7 elsewhere
    ph = g["nut2_colony_pH"]
    assert np.all((ph > 0) & (ph <= 14))


def test_env_params_from_demo_file_equal_defaults():
    from baccolonygrowth_amd.environment import default_env_params, env_params_from_file

    a, b = env_params_from_file(DEMO_INPUT), default_env_params()
    for f, _ in a._fields_:
        assert getattr(a, f) == getattr(b, f), f


def test_clock_flags_follow_double_decrement():
    """bcg_clock_* against the reference's flag rule (ABM/Simulation.cpp:415-426)
    restated in Python doubles."""
    from baccolonygrowth_amd.environment import advance, clock

    dt, out_t, upd_t = 1e-4, 0.01, 1e-3
    c = clock(0.0, dt, out_t, upd_t)
    nu, nh, no = upd_t, upd_t / 10.0, out_t
    fu, fh, fo = True, True, False
    for _ in range(500):
        assert (bool(c.UpdateFlag), bool(c.HeightDensityUpdateFlag), bool(c.OutFlag)) == (fu, fh, fo)
        advance(c)
        fo = no <= 0
        no = (out_t if fo else no) - dt
        fu = nu <= 0
        nu = (upd_t if fu else nu) - dt
        fh = nh <= 0
        nh = (upd_t / 10.0 if fh else nh) - dt
    # the update period that double arithmetic gives: 11 steps, then 10
    c = clock(0.0, dt, out_t, upd_t)
    steps = []
    for s in range(60):
        if c.UpdateFlag:
            steps.append(s)
        advance(c)
    assert np.diff(steps).tolist() == [11, 10, 10, 10, 10]


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libbcgref.so")),
                    reason="oracle/_ref (the reference compiled from source) is not built here")
def test_reference_reproduces_golden_env():
    """The reference's passes are deterministic: a rerun gives the fixture."""
    from oracle.oracle import RefEnv, RefLib

    g = np.load(GOLDEN_ENV)
    p, ep = _params(g, 1e-4)
    ref = RefLib()
    _, _, cells = restart_cells(5)
    sim = ref.sim(p)
    sim.set_cells(cells)
    env = RefEnv(ref, sim, ep)
    env.fields(0)
    f = env.get_fields()
    assert np.array_equal(f["density_shift"], g["init_density_shift"])
    assert env.nutrients() == int(g["n1"])
    st = env.get_state()
    assert np.array_equal(st["colony"]["Carbon"], g["nut1_colony_Carbon"])
    env.close()
    sim.close()


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libbcgref.so")),
                    reason="oracle/_ref (the reference compiled from source) is not built here")
def test_density_and_env_files_match_reference(tmp_path):
    """The output step's density and environment files written from the
    state arrays equal the reference's own writers' files, byte for byte."""
    from baccolonygrowth_amd.environment import write_density_text, write_env_text
    from oracle.oracle import RefEnv, RefLib

    g = np.load(GOLDEN_ENV)
    p, ep = _params(g, 1e-4)
    ref = RefLib()
    _, _, cells = restart_cells(5)
    sim = ref.sim(p)
    sim.set_cells(cells)
    env = RefEnv(ref, sim, ep)
    env.fields(0)
    env.nutrients()
    env.output(str(tmp_path / "ref_density.txt"), str(tmp_path / "ref_env.txt"))
    f, st = env.get_fields(), env.get_state()
    write_density_text(str(tmp_path / "density.txt"), f["density_shift"], ep.BoxX, ep.BoxZ)
    write_env_text(str(tmp_path / "env.txt"), st["colony"], ep.BoxX, ep.BoxZ)
    env.close()
    sim.close()
    for a, b in (("ref_density.txt", "density.txt"), ("ref_env.txt", "env.txt")):
        ra, rb = (tmp_path / a).read_bytes(), (tmp_path / b).read_bytes()
        assert len(ra) > 1000 and ra == rb, a


# ------------------------------------------------------------------ GPU
def _start_fields(col, p):
    """Height = cellRadius, normal = (0, 0, -1): the reference's initial
    arrays (on both sides)."""
    nh = p.BoxX * p.refinementGridHeight * p.BoxY * p.refinementGridHeight
    normal = np.zeros((nh, 3))
    normal[:, 2] = -1.0
    col.set_fields(height=np.full(nh, p.cellRadius), normal=normal)


def _check_fields(env, col, g, prefix):
    f = env.get_fields()
    assert np.array_equal(f["bounds"], g[f"{prefix}_bounds"]), (prefix, f["bounds"], g[f"{prefix}_bounds"])
    for k in ("density", "density_shift", "inside", "wall_shift"):
        _close(f[k], g[f"{prefix}_{k}"], f"{prefix} {k}")
    height, normal, _, _, _ = col.fields()
    _close(height, g[f"{prefix}_height"], f"{prefix} height")
    _close(normal, g[f"{prefix}_normal"], f"{prefix} normal")


def _check_state(env, g, prefix, agar=True):
    st = env.state()
    worst = 0.0
    for k in FIELDS + BUFFER_FIELDS:
        tol = RTOL_BUFFER if k in BUFFER_FIELDS else RTOL
        worst = max(worst, _close(st["colony"][k], g[f"{prefix}_colony_{k}"], f"{prefix} colony {k}", tol))
        worst = max(worst, _close(st["colony_old"][k], g[f"{prefix}_colony_old_{k}"], f"{prefix} colony_old {k}", tol))
    na = len(g[f"{prefix}_wall_CarbonAgar"])
    bx = na // env.p.maxLevelsMG
    for k in ("CarbonAgar", "OxygenAgar", "AcetateAgar", "WasteAgar"):
        for w in ("wall", "wall_old"):
            got, exp = st[w][k], g[f"{prefix}_{w}_{k}"]
            # levels >= 1: the part covered by the finer level is left with an
            # uninitialised x term by the reference (overwritten each outer
            # iteration); compared outside it only
            keep = np.ones(len(exp), bool)
            for lv in range(1, env.p.maxLevelsMG):
                keep[lv * bx + bx // 4 : lv * bx + bx * 3 // 4] = False
            worst = max(worst, _close(got[keep], exp[keep], f"{prefix} {w} {k}"))
        if agar and k != "WasteAgar":
            exp = g[f"{prefix}_agar_{k}"]
            worst = max(worst, _close(st["agar"][k][: len(exp)], exp, f"{prefix} agar {k}"))
    return worst


@pytest.mark.gpu
@pytest.mark.parametrize("lds", ["lds", "global"])
def test_env_passes_match_reference(lds):
    """INIT fields, nutrients, UPDATE and HEIGHT fields, nutrients again (the
    multigrid state carried over); the swept colony boxes in LDS or in
    global memory."""
    from baccolonygrowth_amd.environment import ENV_HEIGHT, ENV_INIT, ENV_UPDATE, Environment
    from baccolonygrowth_amd.simulation import Colony

    g = np.load(GOLDEN_ENV)
    p, ep = _params(g, 1e-4)
    _, _, cells = restart_cells(5)
    say(f"[gpu] env passes ({lds}): {len(cells)} cells, grid {tuple(g['grid'])}")
    with Colony(p) as col:
        col.set_cells(cells)
        _start_fields(col, p)
        with Environment(col, ep) as env:
            if lds == "global":
                env.set_lds_limit(0)
            env.fields(ENV_INIT)
            _check_fields(env, col, g, "init")
            say("    INIT fields ok")
            assert env.nutrients() == int(g["n1"])
            w = _check_state(env, g, "nut1", agar=False)
            say(f"    nutrients 1 ok (max rel diff {w:.1e})")
            env.fields(ENV_UPDATE)
            _check_fields(env, col, g, "update")
            env.fields(ENV_HEIGHT)
            _check_fields(env, col, g, "height")
            say("    UPDATE and HEIGHT fields ok")
            assert env.nutrients() == int(g["n2"])
            w = _check_state(env, g, "nut2")
            say(f"    nutrients 2 ok (max rel diff {w:.1e})")
            # the mechanics grows with the Environment's growth rates
            _, _, _, gr1, _ = col.fields()
            _close(gr1, g["nut2_colony_GrowthRate1"], "mechanics gr1")


@pytest.mark.gpu
def test_one_nested_agar_level_matches_reference():
    """maxLevelsMG = 1 (the reference's default): the outer level is the
    finest, other composite-boundary branches."""
    from baccolonygrowth_amd.environment import ENV_INIT, Environment
    from baccolonygrowth_amd.simulation import Colony

    g = np.load(GOLDEN_ENV)
    p, ep = _params(g, 1e-4, grid="ml1_grid", levels=1)
    _, _, cells = restart_cells(5)
    say(f"[gpu] one nested level: grid {tuple(g['ml1_grid'])}")
    with Colony(p) as col:
        col.set_cells(cells)
        _start_fields(col, p)
        with Environment(col, ep) as env:
            env.fields(ENV_INIT)
            _check_fields(env, col, g, "ml1")
            assert env.nutrients() == int(g["ml1_n"])
            w = _check_state(env, g, "ml1")
            say(f"    ok (max rel diff {w:.1e})")


@pytest.mark.gpu
def test_time_loop_with_fields_matches_reference():
    """25 steps of the time loop (bcg_sim_step): field and nutrient updates
    on the clock's flags, list rebuilds, growth with the nutrient-limited
    rates, divisions."""
    from baccolonygrowth_amd.environment import ENV_INIT, Environment, clock, sim_step
    from baccolonygrowth_amd.simulation import Colony

    g = np.load(GOLDEN_ENV)
    update_time = float(g["loop_update_time"])
    p, ep = _params(g, update_time)
    t0, dt, _ = restart_cells(5)
    cells = np.ascontiguousarray(g["loop_cells_in"]).view(CELL_DTYPE).reshape(-1).copy()
    say(f"[gpu] time loop: {len(cells)} cells, {int(g['loop_steps'])} steps")
    with Colony(p) as col:
        col.set_cells(cells)
        _start_fields(col, p)
        col.set_fields(wall=g["loop_wall"])
        col.srand(int(g["loop_seed"]))
        with Environment(col, ep) as env:
            env.fields(ENV_INIT)
            env.nutrients()
            c = clock(t0, dt, 1.0, update_time)
            lineage = []
            for _ in range(int(g["loop_steps"])):
                lineage += sim_step(col, env, c).tolist()
            assert np.array_equal(np.array(lineage, np.int32).reshape(-1, 2), g["loop_lineage"])
            assert c.t == float(g["loop_t"])
            got = col.cells()
            exp = np.ascontiguousarray(g["loop_cells"]).view(CELL_DTYPE).reshape(-1)
            assert len(got) == len(exp)
            for f in ("p", "q"):
                d = np.abs(got[f] - exp[f]).max()
                assert d <= 1e-10 * max(np.abs(exp[f]).max(), 1.0), (f, d)
            for f in ("Length", "Ldiv", "GrowthRate"):
                d = np.abs(got[f] - exp[f]).max()
                assert d <= 1e-10 * max(np.abs(exp[f]).max(), 1.0), (f, d)
            _check_fields(env, col, g, "loop")
            w = _check_state(env, g, "loop")
            say(f"    loop ok: {len(lineage)} divisions, env max rel diff {w:.1e}")


@pytest.mark.gpu
def test_demo_grid_update_is_bounded_and_placement_independent():
    """The demo input itself (256 x 50 boxes, BoxZAgar = 32: the coarsest
    multigrid level has one z point, and the reference's stencils read past
    its array there - parity unpinned for that read, 0 here) on the demo
    restart state 8: concentrations stay within their bounds, the swept boxes
    in LDS or in global memory give identical states, and a rerun repeats
    them bit for bit."""
    from baccolonygrowth_amd.environment import ENV_INIT, ENV_UPDATE, Environment, env_params_from_file
    from baccolonygrowth_amd.simulation import Colony

    p = demo_params()
    p.maxCells = 5000
    ep = env_params_from_file(DEMO_INPUT)
    ep.UpdateTime = 1e-4
    _, _, cells = restart_cells(8)
    say(f"[gpu] demo grid: {len(cells)} cells, {ep.BoxX}x{ep.BoxZ}, agar {ep.BoxZAgar}")
    states = []
    for lds in (None, 0, None):
        with Colony(p) as col:
            col.set_cells(cells)
            _start_fields(col, p)
            with Environment(col, ep) as env:
                if lds is not None:
                    env.set_lds_limit(lds)
                env.fields(ENV_INIT)
                env.nutrients()
                env.fields(ENV_UPDATE)
                env.nutrients()
                states.append(env.state())
    for k in states[0]:
        a = states[0][k].view(np.uint8)
        assert np.array_equal(a, states[1][k].view(np.uint8)), f"{k}: LDS and global placement differ"
        assert np.array_equal(a, states[2][k].view(np.uint8)), f"{k}: rerun differs"
    c, w = states[0]["colony"], states[0]["wall"]
    assert np.all(np.isfinite(c["Carbon"])) and np.all((c["Carbon"] >= 0) & (c["Carbon"] <= ep.maxCarbon))
    assert np.all((c["Oxygen"] >= 0) & (c["Oxygen"] <= ep.maxO2))
    assert np.all(c["Acetate"] >= 0) and np.all(c["Waste"] >= 0)
    assert np.all((w["CarbonAgar"] >= 0) & (w["CarbonAgar"] <= ep.maxCarbon))
    gr = c["GrowthRate1"]
    assert np.all(np.isfinite(gr)) and gr.max() > 0
    say(f"    ok: growth rate range [{gr.min():.3g}, {gr.max():.3g}]")


# ------------------------------------------------ the reference's large input
GOLDEN_LARGE = os.path.join(GOLDEN, "golden_env_large.npz")
LARGE_INPUT = os.path.join(GOLDEN, "inputFile_20mM_LargeSimulation.txt")
LARGE_STRIDE = (8, 16)  # the fixture's agar samples: every 8th z, every 16th x


def _large_params(g):
    """run_sim/inputFile_20mM_LargeSimulation.txt as the reference reads it,
    with the fixture's Update_Time (2 outer iterations an update)."""
    from baccolonygrowth_amd.cells import read_parameters
    from baccolonygrowth_amd.environment import env_params_from_file

    p, _ = read_parameters(LARGE_INPUT)
    p.maxNeighbours, p.planar, p.maxCells = 96, 1, 4000
    ep = env_params_from_file(LARGE_INPUT)
    ep.UpdateTime = float(g["update_time"])
    return p, ep


def _agar_weights(n, k):
    return np.random.default_rng(1000 + k).random(n)


def _check_large_state(st, g, ep, prefix):
    bx, bz, bza, ml = ep.BoxX, ep.BoxZ, ep.BoxZAgar, ep.maxLevelsMG
    x0, x1 = (int(v) for v in g["window"])
    worst = 0.0
    wc = _agar_weights(bx * bz, 2)
    for k in FIELDS + BUFFER_FIELDS:
        tol = RTOL_BUFFER if k in BUFFER_FIELDS else RTOL
        for b in ("colony", "colony_old"):
            full = st[b][k].reshape(bx, bz)
            worst = max(worst, _close(full[x0:x1], g[f"{prefix}_{b}_{k}"], f"{prefix} {b} {k}", tol))
            ws, asum = g[f"{prefix}_{b}_{k}_sums"]
            assert abs(st[b][k] @ wc - ws) <= 10 * tol * max(asum, 1e-300), (prefix, b, k)
    for k in ("CarbonAgar", "OxygenAgar", "AcetateAgar", "WasteAgar"):
        for w in ("wall", "wall_old"):
            got, exp = st[w][k], g[f"{prefix}_{w}_{k}"]
            keep = np.ones(len(exp), bool)  # as _check_state: the coarser levels' middle is unpinned
            for lv in range(1, ml):
                keep[lv * bx + bx // 4: lv * bx + bx * 3 // 4] = False
            worst = max(worst, _close(got[keep], exp[keep], f"{prefix} {w} {k}"))
        if k == "WasteAgar":
            continue
        a = st["agar"][k].reshape(ml, bx, bza)
        worst = max(worst, _close(a[:, :: LARGE_STRIDE[1], :: LARGE_STRIDE[0]], g[f"{prefix}_agar_{k}_sample"],
                                  f"{prefix} agar {k} samples"))
        sums = g[f"{prefix}_agar_{k}_sums"]
        w0, w1 = _agar_weights(bx * bza, 0), _agar_weights(bx * bza, 1)
        for lv in range(ml):
            got = np.array([a[lv].ravel() @ w0, a[lv].ravel() @ w1])
            # each term within 1e-13 of its value: the sums within 1e-13 of sum |x|
            assert np.all(np.abs(got - sums[lv, :2]) <= 1e-12 * sums[lv, 2]), (prefix, k, lv, got, sums[lv])
    return worst


def test_golden_env_large_fixture_is_consistent():
    g = np.load(GOLDEN_LARGE)
    bx, bz, bza, ml = (int(v) for v in g["grid"])
    assert (bx, bz, bza, ml) == (1520, 100, 512, 3)
    assert bza >= 2 ** (8 - ml)  # every multigrid level keeps z points
    x0, x1 = (int(v) for v in g["window"])
    b = g["init_bounds"]
    assert x0 <= b[0] - 2 and x1 >= b[1] + 3  # the window holds every swept box
    assert int(g["n1"]) == int(3600 * float(g["update_time"]) * 740 / 100) * 200 == int(g["n2"])
    assert g["nut1_colony_Carbon"].shape == (x1 - x0, bz)
    assert g["nut2_agar_CarbonAgar_sample"].shape == (ml, (bx + 15) // 16, (bza + 7) // 8)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libbcgref.so")),
                    reason="oracle/_ref (the reference compiled from source) is not built here")
def test_reference_reproduces_golden_env_large():
    from baccolonygrowth_amd.environment import ENV_DTYPE  # noqa: F401
    from oracle.oracle import RefEnv, RefLib

    g = np.load(GOLDEN_LARGE)
    p, ep = _large_params(g)
    ref = RefLib()
    _, _, cells = restart_cells(int(g["restart"]))
    sim = ref.sim(p)
    sim.set_cells(cells)
    env = RefEnv(ref, sim, ep)
    env.fields(0)
    assert np.array_equal(env.get_fields()["density_shift"], g["init_density_shift"])
    assert env.nutrients() == int(g["n1"])
    _check_large_state(env.get_state(), g, ep, "nut1")
    env.close()
    sim.close()


@pytest.mark.gpu
def test_large_simulation_grid_matches_reference():
    """The reference's second shipped input: 1520 x 100 colony boxes, three
    nested 1521 x 513 agar levels and five coarser ones, on the demo restart
    state 8 - INIT fields, a nutrient update, UPDATE and HEIGHT fields, a
    second update with the multigrid state carried over."""
    from baccolonygrowth_amd.environment import ENV_HEIGHT, ENV_INIT, ENV_UPDATE, Environment
    from baccolonygrowth_amd.simulation import Colony

    g = np.load(GOLDEN_LARGE)
    p, ep = _large_params(g)
    _, _, cells = restart_cells(int(g["restart"]))
    say(f"[gpu] LargeSimulation grid {ep.BoxX}x{ep.BoxZ}, agar {ep.BoxZAgar} x {ep.maxLevelsMG} levels, "
        f"{len(cells)} cells")
    with Colony(p) as col:
        col.set_cells(cells)
        _start_fields(col, p)
        with Environment(col, ep) as env:
            env.fields(ENV_INIT)
            f = env.get_fields()
            assert np.array_equal(f["bounds"], g["init_bounds"])
            for k in ("density_shift", "inside", "wall_shift"):
                _close(f[k], g[f"init_{k}"], f"init {k}")
            assert env.nutrients() == int(g["n1"])
            w = _check_large_state(env.state(), g, ep, "nut1")
            say(f"    update 1 ok (max rel diff {w:.1e})")
            env.fields(ENV_UPDATE)
            env.fields(ENV_HEIGHT)
            assert env.nutrients() == int(g["n2"])
            w = _check_large_state(env.state(), g, ep, "nut2")
            say(f"    update 2 ok (max rel diff {w:.1e})")


@pytest.mark.gpu
@pytest.mark.parametrize("grid", ["demo", "large"])
def test_execution_modes_give_identical_updates(grid):
    """The nutrient update in phases (the default: graph-replayed outer
    iterations, device-wide passes on the wide levels), launched directly,
    with the sweeps in one workgroup or device-wide, and as one workgroup:
    four outer iterations (the recorded graph replayed three times, with
    host synchronisations between the calls), bit for bit the same."""
    from baccolonygrowth_amd.environment import (EXEC_GRID_SWEEPS, EXEC_NO_GRAPH, EXEC_ONE_WORKGROUP, EXEC_WG_SWEEPS,
                                                 ENV_INIT, ENV_UPDATE, Environment, env_params_from_file)
    from baccolonygrowth_amd.simulation import Colony

    if grid == "demo":
        p = demo_params()
        p.maxCells = 5000
        ep = env_params_from_file(DEMO_INPUT)
    else:
        p, ep = _large_params(np.load(GOLDEN_LARGE))
    ep.UpdateTime = 4.5 * 100 / (3600 * 740)  # 4 outer iterations
    _, _, cells = restart_cells(8)
    say(f"[gpu] execution modes on the {grid} grid: {len(cells)} cells")
    states = {}
    for flags in (0, EXEC_NO_GRAPH | EXEC_GRID_SWEEPS, EXEC_WG_SWEEPS, EXEC_ONE_WORKGROUP):
        with Colony(p) as col:
            col.set_cells(cells)
            _start_fields(col, p)
            with Environment(col, ep) as env:
                env.set_execution(flags)
                env.fields(ENV_INIT)
                assert env.nutrients() == 800
                env.fields(ENV_UPDATE)
                env.nutrients()
                states[flags] = env.state()
    ref = states[0]
    for flags, st in states.items():
        for k in ref:
            assert np.array_equal(st[k].view(np.uint8), ref[k].view(np.uint8)), f"flags {flags}: {k} differs"
    say("    all modes equal")


GOLDEN_OUTLOOP = os.path.join(GOLDEN, "golden_output_loop.npz")


@pytest.mark.gpu
def test_time_loop_with_output_steps_matches_reference():
    """The time loop split at its output point (bcg_sim_prepare, then at
    OutFlag the cell file's mean_stress, then bcg_sim_advance with the moved
    cells that the Restart file saves), with output steps falling on update
    steps, against the reference's own loop order: stress and net force of
    every cell, the saved cells, lineage, final cells and the rand()
    position."""
    from baccolonygrowth_amd.environment import ENV_INIT, Environment, clock, sim_advance, sim_prepare
    from baccolonygrowth_amd.simulation import Colony

    g = np.load(GOLDEN_OUTLOOP)
    update_time = float(g["update_time"])
    p, ep = _params(np.load(GOLDEN_ENV), update_time)
    t0, dt, _ = restart_cells(5)
    cells = np.ascontiguousarray(g["cells_in"]).view(CELL_DTYPE).reshape(-1).copy()
    out_steps = [int(k) for k in g["out_steps"]]
    say(f"[gpu] time loop with output steps {out_steps}: {len(cells)} cells")
    with Colony(p) as col:
        col.set_cells(cells)
        _start_fields(col, p)
        col.set_fields(wall=g["wall"])
        col.srand(int(g["seed"]))
        with Environment(col, ep) as env:
            env.fields(ENV_INIT)
            env.nutrients()
            c = clock(t0, dt, float(g["output_time"]), update_time)
            lineage, seen = [], []
            for k in range(int(g["steps"])):
                out = bool(c.OutFlag)
                assert out == (k in out_steps), k
                sim_prepare(col, env, c)
                if out:
                    S, F = col.mean_stress()
                    for name, got, exp in (("stress", S, g[f"out{k}_stress"]), ("fnet", F, g[f"out{k}_fnet"])):
                        err = np.abs(got - exp).max() / np.abs(exp).max()
                        assert err <= 1e-10, (k, name, err)
                lin, moved = sim_advance(col, c, moved=out)
                lineage += lin.tolist()
                if out:
                    exp = np.ascontiguousarray(g[f"out{k}_moved"]).view(CELL_DTYPE).reshape(-1)
                    assert len(moved) == len(exp)
                    for f in ("p", "q", "Length"):
                        d = np.abs(moved[f] - exp[f]).max()
                        assert d <= 1e-10 * max(np.abs(exp[f]).max(), 1.0), (k, f, d)
                    seen.append(k)
            assert seen == out_steps
            assert np.array_equal(np.array(lineage, np.int32).reshape(-1, 2), g["lineage"])
            got = col.cells()
            exp = np.ascontiguousarray(g["cells"]).view(CELL_DTYPE).reshape(-1)
            assert len(got) == len(exp)
            for f in ("p", "q"):
                d = np.abs(got[f] - exp[f]).max()
                assert d <= 1e-10 * max(np.abs(exp[f]).max(), 1.0), (f, d)
            assert [col.rand() for _ in range(4)] == g["next_draws"].tolist()
            say(f"    ok: {len(lineage)} divisions, output steps {seen}")
