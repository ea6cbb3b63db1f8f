"""Parity of the HIP mechanics path (through the C ABI) with the reference.

Checked against the reference's golden vectors (tests/golden, written by the
reference's own code) and against the oracle restatement on the same inputs:
  * neighbour sets: bit-exact;
  * division events (lineage) and the rand() stream position: exact;
  * forces: |dF| <= 1e-10 * (|F| + sum_j |F_ij|) per component;
  * positions: within 1e-10 relative, velocities within 1e-8 of the
    velocity scale (FP64 throughout; summation order differs).
"""
import os
import sys

import numpy as np
import pytest

from baccolonygrowth_amd.cells import CELL_DTYPE
from helpers import GOLDEN, demo_params, planar_case, sorted_lists

pytestmark = pytest.mark.gpu

POS_RTOL = 1e-10
FORCE_RTOL = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def Colony():
    from baccolonygrowth_amd.simulation import Colony

    return Colony


def _cells(a):
    return np.ascontiguousarray(a).view(CELL_DTYPE).reshape(-1).copy()


def _abs_pair_sums(oracle, p, cells, counts, lists, idx=None):
    """sum_j |F_ij| per cell (scale of the cancellation in sum_forces)."""
    S = np.zeros(len(cells))
    for i in range(len(cells)) if idx is None else idx:
        for j in lists[i, : counts[i]]:
            f = oracle.F_cc(p, cells[i : i + 1], cells[j : j + 1])
            S[i] += np.abs(f[:3]).max()
    return S


def _check_cells(a, b, label):
    assert len(a) == len(b), label
    for f in ("age_p", "age_q", "Type", "Ancestor"):
        assert np.array_equal(a[f], b[f]), (label, f)
    for f in ("Length", "Ldiv", "Radius", "GrowthRate", "time_p", "time_q"):
        d = np.abs(a[f] - b[f]).max()
        assert d <= 1e-12 * max(np.abs(b[f]).max(), 1.0), (label, f, d)
    for f in ("p", "q"):
        d = np.abs(a[f] - b[f]).max()
        assert d <= POS_RTOL * max(np.abs(b[f]).max(), 1.0), (label, f, d)
    for f in ("Velocity", "AngularVelocity", "DynFric", "StaFric"):
        d = np.abs(a[f] - b[f]).max()
        assert d <= 1e-8 * max(np.abs(b[f]).max(), 1.0), (label, f, d)


@pytest.mark.parametrize("k", [7, 8])
def test_neighbours_and_forces_vs_reference_golden(Colony, oracle, k):
    z = np.load(os.path.join(GOLDEN, f"golden_step_r{k}.npz"))
    p = demo_params()
    cells = _cells(z["cells_in"])
    say(f"[gpu] golden r{k}: {len(cells)} cells")
    with Colony(p) as col:
        col.LoadCells(cells)
        col.set_fields(z["height"], z["normal"], z["wall"], z["gr1"], z["gr2"])
        col.getNeighbours()
        counts, lists = col.neighbours()
        assert np.array_equal(counts, z["nl_counts0"])
        assert sorted_lists(counts, lists) == sorted_lists(z["nl_counts0"], z["nl_lists0"])
        F, T = col.sum_forces()
    S = _abs_pair_sums(oracle, p, cells, counts, lists)
    for got, ref in ((F, z["F0"]), (T, z["T0"])):
        scale = np.abs(ref) + S[:, None] * (1.0 if got is F else 10.0) + 1e-300
        err = np.abs(got - ref) / scale
        say(f"    max rel force/torque error {err.max():.3e}")
        assert err.max() <= FORCE_RTOL


@pytest.mark.parametrize("k", [7, 8])
def test_steps_vs_reference_golden(Colony, k):
    z = np.load(os.path.join(GOLDEN, f"golden_step_r{k}.npz"))
    p = demo_params()
    cells = _cells(z["cells_in"])
    say(f"[gpu] golden r{k}: {int(z['steps'])} steps")
    with Colony(p) as col:
        col.LoadCells(cells)
        col.set_fields(z["height"], z["normal"], z["wall"], z["gr1"], z["gr2"])
        col.getNeighbours()
        col.srand(int(z["seed"]))
        t = float(z["t"])
        lin = []
        for s in range(int(z["steps"])):
            L = col.step(float(z["dt"]), t, update=(s == int(z["rebuild_step"])))
            lin.append(np.column_stack([np.full(len(L), s), L]))
            t += float(z["dt"])
        lin = np.concatenate(lin)
        assert np.array_equal(lin, z["lineage"])
        out = col.cells()
        _check_cells(out, _cells(z["cells_out"]), f"r{k}")
        counts, lists = col.neighbours()
        assert sorted_lists(counts, lists) == sorted_lists(z["nl_counts1"], z["nl_lists1"])
        assert [col.rand() for _ in range(4)] == z["next_draws"].tolist()


def test_many_steps_vs_oracle_planar(Colony, oracle):
    """30 steps with divisions every few steps and list rebuilds on the
    reference cadence: lineage, lists and rand() position exact."""
    p, t, dt, cells, fields = planar_case(7, seed=21, divide_frac=0.15)
    say(f"[gpu] planar 30 steps, {len(cells)} cells")
    o = oracle.sim(p)
    o.set_cells(cells)
    o.set_fields(*fields)
    o.rebuild()
    oracle.srand(5)
    with Colony(p) as col:
        col.LoadCells(cells)
        col.set_fields(*fields)
        col.getNeighbours()
        col.srand(5)
        tt = t
        for s in range(30):
            upd = s % 10 == 0
            la = o.step(dt, tt, rebuild=upd)
            lb = col.step(dt, tt, update=upd)
            assert np.array_equal(la, lb), s
            tt += dt
        _check_cells(col.cells(), o.cells(), "30 steps")
        ca, la_ = o.neighbours(stride=p.maxNeighbours)
        cb, lb_ = col.neighbours()
        assert sorted_lists(ca, la_) == sorted_lists(cb, lb_)
        assert col.rand() == oracle.rand()
    o.close()


def _dome_case(n=20000, seed=2):
    from baccolonygrowth_amd.colony import dome_colony, dome_params, synthetic_fields

    cells = dome_colony(n, radius=110.0 * (n / 20000) ** 0.5, height=12.0, seed=seed)
    p = dome_params(cells, bin_size=5.0, max_neighbours=96, extra_cells=n)
    rng = np.random.default_rng(seed)
    cells["Ldiv"] = cells["Length"] * (1.0 + rng.uniform(0, 3e-3, n))
    cells["Velocity"] = rng.normal(0, 0.5, (n, 3))
    cells["AngularVelocity"] = rng.normal(0, 0.5, (n, 3))
    fields = synthetic_fields(p, cells, seed=seed)
    return p, cells, fields


def test_3d_dome_vs_oracle(Colony, oracle):
    """3-D model: lists exact, forces at 1e-10, 3 steps with divisions."""
    p, cells, fields = _dome_case()
    say(f"[gpu] 3-D dome {len(cells)} cells, bins {p.BoxX}x{p.BoxY}x{p.BoxZ}")
    o = oracle.sim(p)
    o.set_cells(cells)
    o.set_fields(*fields)
    o.rebuild()
    with Colony(p) as col:
        col.LoadCells(cells)
        col.set_fields(*fields)
        col.getNeighbours()
        ca, la = o.neighbours(stride=p.maxNeighbours)
        cb, lb = col.neighbours()
        assert np.array_equal(ca, cb)
        assert sorted_lists(ca, la) == sorted_lists(cb, lb)
        Fa, Ta = o.sum_forces()
        Fb, Tb = col.sum_forces()
        sel = np.arange(0, len(cells), 37)
        S = _abs_pair_sums(oracle, p, cells, ca, la, sel)
        err = np.abs(Fb - Fa)[sel] / (np.abs(Fa)[sel] + S[sel, None] + 1e-300)
        say(f"    3-D max rel force error {err.max():.3e}")
        assert err.max() <= FORCE_RTOL
        oracle.srand(11)
        col.srand(11)
        tt = 0.0
        dt = 1.0 / 3600
        nd = 0
        for s in range(3):
            la_ = o.step(dt, tt, rebuild=(s == 1))
            lb_ = col.step(dt, tt, update=(s == 1))
            assert np.array_equal(la_, lb_)
            nd += len(la_)
            tt += dt
        assert nd > 0
        _check_cells(col.cells(), o.cells(), "3-D")
        assert col.rand() == oracle.rand()
    o.close()


def test_rebuild_every_step_no_patch_equals_patch(Colony):
    """Skipping the post-division patch when the next step rebuilds gives
    bit-identical cells (the patched lists are never read)."""
    p, cells, fields = _dome_case(8000, seed=5)
    res = []
    for no_patch in (False, True):
        with Colony(p) as col:
            col.LoadCells(cells)
            col.set_fields(*fields)
            col.srand(3)
            for s in range(4):
                col.step(1.0 / 3600, s / 3600.0, update=True, no_patch=no_patch)
            res.append(col.cells())
    assert res[0].tobytes() == res[1].tobytes()


def test_deterministic_bitwise(Colony):
    p, cells, fields = _dome_case(8000, seed=6)
    outs = []
    for _ in range(2):
        with Colony(p) as col:
            col.LoadCells(cells)
            col.set_fields(*fields)
            col.srand(9)
            for s in range(3):
                col.step(1.0 / 3600, s / 3600.0, update=True)
            outs.append((col.cells().tobytes(), col.neighbours()[0].tobytes()))
    assert outs[0] == outs[1]


# ----------------------------------------------------------- edge cases ---
def test_empty_and_single(Colony):
    p = demo_params()
    with Colony(p) as col:
        col.LoadCells(np.zeros(0, CELL_DTYPE))
        assert col.step(1e-4, 0.0, update=True).shape == (0, 2)
        assert col.n == 0
        c = np.zeros(1, CELL_DTYPE)
        c["p"] = (1.0, 0, 0.5)
        c["q"] = (-1.0, 0, 0.5)
        c["Length"] = 2.0
        c["Radius"] = 0.5
        c["Ldiv"] = 10.0
        c["Type"] = 1
        c["Ancestor"] = 1
        col.LoadCells(c)
        col.step(1e-4, 0.0, update=True)
        assert col.n == 1
        assert col.neighbours()[0].tolist() == [0]


def test_out_of_domain_is_an_error(Colony):
    from baccolonygrowth_amd._lib import BcgError

    p = demo_params()
    c = np.zeros(1, CELL_DTYPE)
    c["p"] = (1e4, 0, 0.5)
    c["q"] = (1e4 - 2, 0, 0.5)
    c["Length"] = 2.0
    c["Radius"] = 0.5
    with Colony(p) as col:
        with pytest.raises(BcgError) as e:
            col.LoadCells(c)
        assert e.value.code == -3


def test_failed_step_is_sticky_until_set_cells(Colony):
    """A step that fails after its move has run (a cell pushed across the
    +x edge of the demo grid, 512 um, by a neighbour 1.2 um behind it: about
    2.7e-4 um in one step) leaves the context failed: every later call that
    uses its cells reports BCG_ERR_STATE until bcg_set_cells, after which it
    steps normally."""
    from baccolonygrowth_amd._lib import BcgError

    p = demo_params()
    p.maxCells = 16
    edge = p.BoxX / 2 * p.BoxLength
    c = np.zeros(2, CELL_DTYPE)
    for i, cx in enumerate((edge - 1e-5, edge - 1e-5 - 1.2)):
        c["p"][i] = (cx + 0.5, 0, 0.5)
        c["q"][i] = (cx - 0.5, 0, 0.5)
    c["Length"], c["Radius"], c["Ldiv"], c["Type"], c["Ancestor"] = 1.0, 0.5, 10.0, 1, 1
    nh = p.BoxX * p.refinementGridHeight
    normal = np.tile([0.0, 0.0, -1.0], nh)
    with Colony(p) as col:
        col.LoadCells(c)
        col.set_fields(height=np.full(nh, 0.5), normal=normal, wall=np.zeros(p.BoxX))
        with pytest.raises(BcgError) as e:
            col.step(1e-4, 0.0, update=True)
        assert e.value.code == -3
        for call in (col.cells, lambda: col.step(1e-4, 0.0, update=True), col.getNeighbours, col.mean_stress,
                     col.sum_forces):
            with pytest.raises(BcgError) as e:
                call()
            assert e.value.code == -7
        safe = c.copy()
        safe["p"][:, 0] -= 100.0
        safe["q"][:, 0] -= 100.0
        col.LoadCells(safe)
        col.step(1e-4, 0.0, update=True)
        assert col.n == 2 and np.all(np.isfinite(col.cells()["p"]))


def test_neighbour_overflow_is_an_error(Colony):
    from baccolonygrowth_amd._lib import BcgError

    p, t, dt, cells, fields = planar_case(8, seed=1)
    p = p.replace(maxNeighbours=2)
    with Colony(p) as col:
        col.LoadCells(cells)
        with pytest.raises(BcgError) as e:
            col.getNeighbours()
        assert e.value.code == -4


def test_max_cells_stops_division(Colony, oracle):
    """Reaching MaxCells stops the division loop (Simulation.cpp:382-386)."""
    p, t, dt, cells, fields = planar_case(5, seed=8, divide_frac=0.5)
    p = p.replace(maxCells=len(cells) + 3)
    o = oracle.sim(p)
    o.set_cells(cells)
    o.set_fields(*fields)
    o.rebuild()
    oracle.srand(4)
    with Colony(p) as col:
        col.LoadCells(cells)
        col.set_fields(*fields)
        col.getNeighbours()
        col.srand(4)
        for s in range(3):
            la = o.step(dt, t + s * dt, rebuild=False)
            lb = col.step(dt, t + s * dt, update=False)
            assert np.array_equal(la, lb)
        assert col.n == len(cells) + 3 == o.n
        _check_cells(col.cells(), o.cells(), "maxCells")
    o.close()


def test_partial_rebuild_matches_oracle(Colony, oracle):
    p, t, dt, cells, fields = planar_case(8, seed=2)
    ids = np.arange(0, len(cells), 13)
    o = oracle.sim(p)
    o.set_cells(cells)
    o.rebuild(ids)
    with Colony(p) as col:
        col.LoadCells(cells)
        col.getNeighbours(ids)
        ca, la = o.neighbours(stride=p.maxNeighbours)
        cb, lb = col.neighbours()
        assert sorted_lists(ca, la) == sorted_lists(cb, lb)
    o.close()


def _threshold_case(seed=7):
    """Pairs of cells, 20 um apart from every other pair and several hundred um
    from the origin, each placed at one of the reference predicate's
    thresholds (ABM/Neighbours.cpp:72-88), to relative offsets from 2e-16 to
    1e-6 on either side:
      * B crossing A at 0.8 um (the second test passes) with both end points at
        distance (L_A/2 + Rc R_A)(1 + d) from the centre of A (the first test
        decides);
      * B nearly parallel to A at a distance Rc R_A (1 + d) (the second test
        decides, the first one passing).
    The FP32 candidate prefilter must keep every pair the FP64 predicate keeps,
    and the square roots must be decided as the reference decides them."""
    from baccolonygrowth_amd.cells import reference_defaults
    from baccolonygrowth_amd.colony import dome_params

    rng = np.random.default_rng(seed)
    Rc = reference_defaults().Rc
    R = 0.5
    offsets = [s * m for m in (2e-16, 1e-15, 4e-15, 1e-14, 1e-12, 1e-9, 1e-6) for s in (-1.0, 1.0)]
    cells = []

    def rod(c, u, L):
        return (c - 0.5 * L * u, c + 0.5 * L * u, L)

    def unit(v):
        return v / np.linalg.norm(v)

    k = 0
    for family in ("crossing", "parallel"):
        for d in offsets:
            for rep in range(6):
                c = np.array([420.0 + 20.0 * (k % 12), -640.0 + 20.0 * (k // 12), 1.5 + rng.uniform(0.0, 3.0)])
                k += 1
                uA = unit(np.array([np.cos(a := rng.uniform(0, 2 * np.pi)), np.sin(a), rng.normal(0, 0.15)]))
                LA = rng.uniform(2.0, 4.0)
                pA, qA, _ = rod(c, uA, LA)
                LB = rng.uniform(2.0, 4.0)
                if family == "crossing":
                    n = unit(np.cross(uA, rng.normal(size=3)))
                    uB = np.cross(uA, n)
                    h = 0.8
                    half = np.sqrt(((LA / 2 + Rc * R) * (1.0 + d)) ** 2 - h * h)
                    pB, qB = c + n * h - uB * half, c + n * h + uB * half
                else:
                    n = unit(np.cross(uA, rng.normal(size=3)))
                    uB = unit(uA + 1e-3 * unit(np.cross(uA, n)))
                    cB = c + n * (Rc * R) * (1.0 + d) + uA * rng.uniform(-0.3, 0.3)
                    pB, qB, _ = rod(cB, uB, LB)
                for pp, qq, LL in ((pA, qA, LA), (pB, qB, np.linalg.norm(qB - pB))):
                    cells.append((pp, qq, LL))
    a = np.zeros(len(cells), CELL_DTYPE)
    for i, (pp, qq, LL) in enumerate(cells):
        a[i]["p"] = pp
        a[i]["q"] = qq
        a[i]["Length"] = LL
        a[i]["Radius"] = R
        a[i]["Ldiv"] = 10.0
        a[i]["Type"] = 1
    p = dome_params(a, bin_size=5.0, max_neighbours=16, extra_cells=64)
    return p, a


def test_lists_at_the_predicate_thresholds(Colony, oracle):
    """Neighbour sets equal the oracle's on pairs placed at the predicate's
    thresholds far from the origin (FP32 prefilter margin, square roots
    decided near the threshold)."""
    p, cells = _threshold_case()
    say(f"[gpu] threshold pairs: {len(cells) // 2} pairs, bins {p.BoxX}x{p.BoxY}x{p.BoxZ}")
    o = oracle.sim(p)
    o.set_cells(cells)
    o.rebuild()
    with Colony(p) as col:
        col.LoadCells(cells)
        col.getNeighbours()
        ca, la = o.neighbours(stride=p.maxNeighbours)
        cb, lb = col.neighbours()
        assert np.array_equal(ca, cb)
        assert sorted_lists(ca, la) == sorted_lists(cb, lb)
    o.close()
    # the case decides both ways at both thresholds
    kept = ca.reshape(2, -1, 2)[:, :, 0]  # anchor's list length, per family and pair
    say(f"    kept: crossing family {int((kept[0] > 0).sum())}/{kept.shape[1]}, parallel family {int((kept[1] > 0).sum())}/{kept.shape[1]}")
    for fam in kept:
        assert 0 < (fam > 0).sum() < len(fam)


@pytest.mark.parametrize("case", ["planar", "3-D"])
def test_mean_stress_vs_oracle(Colony, oracle, case):
    """mean_stress of every cell (the output step's stress tensor and net
    force) against the oracle, on a planar state and on the 3-D dome; the
    F_cw draws leave both streams at the same position."""
    if case == "planar":
        p, t, dt, cells, fields = planar_case(5, seed=56, divide_frac=0.2)
    else:
        p, cells, fields = _dome_case(n=8000, seed=3)
    say(f"[gpu] mean_stress {case}, {len(cells)} cells")
    o = oracle.sim(p)
    o.set_cells(cells)
    o.set_fields(*fields)
    o.rebuild()
    oracle.srand(7)
    Sa, Fa = o.mean_stress()
    ra = oracle.rand()
    o.close()
    with Colony(p) as col:
        col.LoadCells(cells)
        col.set_fields(*fields)
        col.getNeighbours()
        col.srand(7)
        Sb, Fb = col.mean_stress()
        rb = col.rand()
    assert ra == rb
    for name, a, b in (("stress", Sa, Sb), ("Fnet", Fa, Fb)):
        err = np.abs(b - a).max() / np.abs(a).max()
        say(f"    {case} {name} max rel error {err:.2e}")
        assert err <= 1e-10


def _sorted_rows(counts, lists):
    """Each row's first counts[i] entries sorted ascending, the rest padded
    (set comparison of all lists at once)."""
    m = np.arange(lists.shape[1])[None, :] < counts[:, None]
    out = np.where(m, lists, np.iinfo(np.int32).max)
    out.sort(axis=1)
    return out


@pytest.mark.timeout(900)
def test_full_size_dome_vs_oracle(Colony, oracle):
    """The 2^20-cell kernel-bench colony itself: neighbour sets of every cell
    exact, then two steps with list rebuilds and divisions (lineage exact,
    positions 1e-10 relative, the rand() position) against the C oracle."""
    from baccolonygrowth_amd.colony import bench_dome

    say("[gpu] full-size dome: building 2^20 cells (about a minute)")
    p, cells, fields = bench_dome(1 << 20)
    say(f"    {len(cells)} cells, bins {p.BoxX}x{p.BoxY}x{p.BoxZ}")
    o = oracle.sim(p)
    o.set_cells(cells)
    o.set_fields(*fields)
    o.rebuild()
    oracle.srand(2)
    with Colony(p) as col:
        col.LoadCells(cells)
        col.set_fields(*fields)
        col.srand(2)
        col.getNeighbours()
        ca, la = o.neighbours(stride=p.maxNeighbours)
        cb, lb = col.neighbours()
        assert np.array_equal(ca, cb)
        assert np.array_equal(_sorted_rows(ca, la), _sorted_rows(cb, lb))
        say(f"    lists equal: mean length {ca.mean():.3f}, max {ca.max()}")
        t, dt, nd = 0.0, 1.0 / 3600, 0
        for s in range(2):
            la_ = o.step(dt, t, rebuild=True)
            lb_ = col.step(dt, t, update=True)
            assert np.array_equal(la_, lb_), f"step {s}: division events differ"
            nd += len(lb_)
            t += dt
            say(f"    step {s}: {len(lb_)} divisions")
        assert nd > 0
        _check_cells(col.cells(), o.cells(), "2^20 dome")
        assert col.rand() == oracle.rand()
    o.close()
