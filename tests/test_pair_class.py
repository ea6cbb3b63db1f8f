"""The FP32 pair decisions of pair_class.h, compiled for the host by
tests/pairs/Makefile: the list build's classification against the oracle's
restatement of reduceNeighbours' test (ABM/Neighbours.cpp:72-88), and the
force gather's "surely not in contact" against the oracle's min_distance
(ABM/Neighbours.cpp:107-160).

Whenever the classifier decides a pair it must decide it as the FP64 predicate
does; the pairs it leaves open go to the FP64 predicate in the kernel. Checked
with the exact reciprocal, with reciprocals one ulp off either way (the kernel
uses the approximate v_rcp_f32), and with one 30 % low: its closest-point
parameters are then far off, and the distance bounds must still hold.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import pytest

from baccolonygrowth_amd.cells import CELL_DTYPE
from oracle.oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
PAIRS = os.path.join(HERE, "pairs")
LIB = os.path.join(PAIRS, "libpairclass.so")
RC, RADIUS, BL = 3.0, 0.5, 5.0


@pytest.fixture(scope="module")
def harness():
    subprocess.run(["make", "-s", "-C", PAIRS], check=True)
    lib = ctypes.CDLL(LIB)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pc_classify_batch.argtypes = [ctypes.c_int, fp, fp, fp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
    lib.pc_apart_batch.argtypes = [ctypes.c_int, fp, fp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
    dp = ctypes.POINTER(ctypes.c_double)
    lib.pc_prefilter_batch.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.POINTER(ctypes.c_int32)]
    return lib


@pytest.fixture(scope="module")
def oracle():
    return Oracle()


def cells_of(c, u, L):
    """Cells with centre c, unit axis u, length L (p, q = c -/+ L/2 u)."""
    out = np.zeros(len(c), CELL_DTYPE)
    out["p"] = c - 0.5 * L[:, None] * u
    out["q"] = c + 0.5 * L[:, None] * u
    out["Length"] = L
    out["Radius"] = RADIUS
    return out


def column_of(pos, nx=456, ny=456):
    """Bin column (x, y) and its origin, as the list build files a centre
    (ABM/UniformGrid.h:87-100 with 5 um bins)."""
    x = np.floor(pos[:, 0] / BL + nx // 2).astype(np.int64)
    y = np.floor(pos[:, 1] / BL + ny // 2 + 0.5).astype(np.int64)
    return x, y, (x - nx // 2) * BL, (y - ny // 2) * BL


def classify(lib, a, b, mode):
    """The classifier on pairs (a[i], b[i]) with FP32 records made as the
    kernels make them: end points relative to the cell's own column origin."""
    ca = 0.5 * (a["p"] + a["q"])
    cb = 0.5 * (b["p"] + b["q"])
    xa, ya, oxa, oya = column_of(ca)
    xb, yb, oxb, oyb = column_of(cb)
    assert np.all(np.abs(xb - xa) <= 1) and np.all(np.abs(yb - ya) <= 1)

    def rec(cells, ox, oy):
        r = np.zeros((len(cells), 2, 4), np.float32)
        for k, e in enumerate(("p", "q")):
            r[:, k, 0] = (cells[e][:, 0] - ox).astype(np.float32)
            r[:, k, 1] = (cells[e][:, 1] - oy).astype(np.float32)
            r[:, k, 2] = cells[e][:, 2].astype(np.float32)
        return r.reshape(len(cells), 8)

    recs = np.ascontiguousarray(np.concatenate([rec(a, oxa, oya), rec(b, oxb, oyb)], axis=1))
    thr = np.ascontiguousarray(
        np.stack([(a["Length"] / 2.0 + RC * a["Radius"]).astype(np.float32),
                  (RC * a["Radius"]).astype(np.float32)], axis=1))
    shift = np.ascontiguousarray(np.stack([((xb - xa) * BL), ((yb - ya) * BL)], axis=1).astype(np.float32))
    out = np.zeros(len(a), np.int32)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pc_classify_batch(len(a), recs.ctypes.data_as(fp), thr.ctypes.data_as(fp), shift.ctypes.data_as(fp), mode,
                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out


def unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def random_axes(rng, n, tilt_deg=15.0):
    phi = rng.uniform(0, 2 * np.pi, n)
    th = np.deg2rad(rng.normal(0, tilt_deg, n))
    return np.stack([np.cos(phi) * np.cos(th), np.sin(phi) * np.cos(th), np.sin(th)], axis=1)


def dome_pairs(rng, n, radius=600.0, height=20.0, reach=6.5):
    """A in a dome-like colony, B with its centre within `reach` of A's and
    in the 27-box (bin columns within one of A's)."""
    r = radius * np.sqrt(rng.random(n))
    ph = rng.uniform(0, 2 * np.pi, n)
    ca = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(0.5, height, n)], axis=1)
    La, Lb = rng.uniform(2, 4, n), rng.uniform(2, 4, n)
    a = cells_of(ca, random_axes(rng, n), La)
    d = unit(rng.normal(size=(n, 3))) * (reach * rng.random(n) ** (1 / 3))[:, None]
    b = cells_of(ca + d, random_axes(rng, n), Lb)
    return keep_in_block(a, b)


def keep_in_block(a, b):
    ca, cb = 0.5 * (a["p"] + a["q"]), 0.5 * (b["p"] + b["q"])
    xa, ya, _, _ = column_of(ca)
    xb, yb, _, _ = column_of(cb)
    ok = (np.abs(xb - xa) <= 1) & (np.abs(yb - ya) <= 1)
    return a[ok], b[ok]


def at_distance(a, b, target):
    """Move b along the common normal of its closest points to a so that the
    FP64 segment distance becomes `target` (pairs whose closest points
    coincide are dropped)."""
    o = Oracle()
    keep, moved = [], b.copy()
    for i in range(len(a)):
        r = o.min_distance(a[i:i + 1], b[i:i + 1])
        d, c1, c2 = r[0], r[1:4], r[4:7]
        if d < 1e-6:
            continue
        nrm = (c2 - c1) / d
        shift = nrm * (target[i] - d)
        moved["p"][i] += shift
        moved["q"][i] += shift
        keep.append(i)
    keep = np.array(keep, np.int64)
    return keep_in_block(a[keep], moved[keep])


def check(lib, oracle, a, b, max_open=None):
    ref = oracle.keep_batch(a, b, RC)
    for mode in (3, 1, 2, 0):
        got = classify(lib, a, b, mode)
        decided = got >= 0
        bad = np.nonzero(decided & (got != ref))[0]
        assert len(bad) == 0, f"mode {mode}: {len(bad)} pairs decided against the FP64 predicate, e.g. {bad[:5]}"
        if max_open is not None and mode != 3:
            assert (~decided).mean() <= max_open, f"mode {mode}: {(~decided).mean():.3%} left open"
    return ref, got


def test_dome_pairs_decided_as_the_fp64_predicate(harness, oracle):
    rng = np.random.default_rng(11)
    a, b = dome_pairs(rng, 200_000)
    ref, got = check(harness, oracle, a, b, max_open=0.002)
    # both verdicts occur in bulk
    assert ref.sum() > 10_000 and (1 - ref).sum() > 10_000


def test_far_from_the_origin_and_tall(harness, oracle):
    rng = np.random.default_rng(12)
    a, b = dome_pairs(rng, 50_000, radius=1200.0, height=60.0)
    check(harness, oracle, a, b, max_open=0.01)


@pytest.mark.parametrize("rel", [1e-2, 1e-4, 1e-7, 1e-10])
def test_segment_distance_at_the_threshold(harness, oracle, rel):
    """Segment distances at Rc R (1 +- rel): decided only where the margin
    allows, and then correctly."""
    rng = np.random.default_rng(int(-np.log10(rel)))
    a, b = dome_pairs(rng, 4000, reach=4.0)
    target = RC * RADIUS * (1.0 + rel * rng.choice([-1.0, 1.0], len(a)))
    a, b = at_distance(a, b, target)
    ref, got = check(harness, oracle, a, b)
    if rel >= 1e-2:
        # the margin is far below 1 %: the classifier decides (almost) all of them
        assert (got < 0).mean() < 0.05


def test_near_parallel_pairs(harness, oracle):
    """B's axis within 1e-3 .. 1e-8 rad of A's (FP32 closest-point parameters
    are ill-determined), at distances around the threshold."""
    rng = np.random.default_rng(21)
    n = 20_000
    a, _ = dome_pairs(rng, n)
    n = len(a)
    ua = unit(a["q"] - a["p"])
    ang = 10.0 ** rng.uniform(-8, -3, n)
    perp = unit(np.cross(ua, rng.normal(size=(n, 3))))
    ub = unit(ua + ang[:, None] * perp)
    side = unit(np.cross(ua, rng.normal(size=(n, 3))))
    dist = RC * RADIUS * rng.uniform(0.5, 1.5, n)
    along = rng.uniform(-3.0, 3.0, n)
    ca = 0.5 * (a["p"] + a["q"])
    cb = ca + dist[:, None] * side + along[:, None] * ua
    b = cells_of(cb, ub, rng.uniform(2, 4, n))
    a, b = keep_in_block(a, b)
    ref, got = check(harness, oracle, a, b, max_open=0.01)


def test_end_point_reach_at_the_threshold(harness, oracle):
    """An end point of B at L/2 + Rc R (1 +- rel) from A's centre, with the
    segment distance well inside: the first test decides."""
    rng = np.random.default_rng(31)
    n = 20_000
    a, _ = dome_pairs(rng, n)
    n = len(a)
    ca = 0.5 * (a["p"] + a["q"])
    reach = a["Length"] / 2.0 + RC * a["Radius"]
    rel = 10.0 ** rng.uniform(-12, -2, n) * rng.choice([-1.0, 1.0], n)
    dirn = unit(rng.normal(size=(n, 3)))
    p = ca + (reach * (1 + rel))[:, None] * dirn
    ub = unit(rng.normal(size=(n, 3)))
    Lb = rng.uniform(2, 4, n)
    b = np.zeros(n, CELL_DTYPE)
    b["p"] = p
    b["q"] = p + Lb[:, None] * ub
    b["Length"] = Lb
    b["Radius"] = RADIUS
    a, b = keep_in_block(a, b)
    check(harness, oracle, a, b)


def test_crossing_and_touching_segments(harness, oracle):
    """Segments that intersect (distance 0) or share an end point."""
    rng = np.random.default_rng(41)
    n = 5000
    a, _ = dome_pairs(rng, n)
    n = len(a)
    s = rng.random(n)
    x = a["p"] + s[:, None] * (a["q"] - a["p"])
    ub = unit(rng.normal(size=(n, 3)))
    Lb = rng.uniform(2, 4, n)
    t = rng.random(n)
    t[: n // 4] = 0.0  # B starts on A
    b = np.zeros(n, CELL_DTYPE)
    b["p"] = x - (t * Lb)[:, None] * ub
    b["q"] = b["p"] + Lb[:, None] * ub
    b["Length"] = Lb
    b["Radius"] = RADIUS
    a, b = keep_in_block(a, b)
    check(harness, oracle, a, b, max_open=0.01)


def test_a_planted_sign_error_is_caught(harness, oracle):
    """The checks above catch a wrong predicate: flipping the oracle's verdict
    on a few decided pairs makes the comparison fail."""
    rng = np.random.default_rng(51)
    a, b = dome_pairs(rng, 5000)
    ref = oracle.keep_batch(a, b, RC)
    got = classify(harness, a, b, 0)
    decided = np.nonzero(got >= 0)[0]
    wrong = ref.copy()
    wrong[decided[:3]] ^= 1
    assert np.any((got >= 0) & (got != wrong))


# ------------------------------------------------------------ contact ---
def apart(lib, a, b, mode):
    """surely_apart on pairs (a[i], b[i]) with end points relative to a's
    centre, as the force gather forms them (FP64 differences, then FP32)."""
    o = 0.5 * (a["p"] + a["q"])
    recs = np.zeros((len(a), 16), np.float32)
    for k, (cells, e) in enumerate(((a, "p"), (a, "q"), (b, "p"), (b, "q"))):
        recs[:, 4 * k:4 * k + 3] = (cells[e] - o).astype(np.float32)
    sigma = np.ascontiguousarray((a["Radius"] + b["Radius"]).astype(np.float32))
    out = np.zeros(len(a), np.int32)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.pc_apart_batch(len(a), np.ascontiguousarray(recs).ctypes.data_as(fp), sigma.ctypes.data_as(fp), mode,
                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out


def oracle_distance(a, b):
    o = Oracle()
    return np.array([o.min_distance(a[k:k + 1], b[k:k + 1])[0] for k in range(len(a))])


def check_apart(lib, a, b, min_skipped=None):
    d = oracle_distance(a, b)
    sigma = a["Radius"] + b["Radius"]
    for mode in (3, 1, 2, 0):
        got = apart(lib, a, b, mode)
        bad = np.nonzero((got == 1) & (d < sigma))[0]
        assert len(bad) == 0, f"mode {mode}: {len(bad)} touching pairs called apart, e.g. d={d[bad[:3]]}"
        if min_skipped is not None and mode != 3:
            frac = got[d >= sigma * 1.01].mean()
            assert frac >= min_skipped, f"mode {mode}: only {frac:.3f} of the clearly separate pairs called apart"
    return d


def test_contact_test_on_dome_neighbours(harness):
    """Neighbour-like pairs (within Rc R): no touching pair is called apart,
    and almost every clearly separate one is."""
    rng = np.random.default_rng(61)
    a, b = dome_pairs(rng, 30_000, reach=3.0)
    check_apart(harness, a, b, min_skipped=0.99)


@pytest.mark.parametrize("rel", [1e-2, 1e-5, 1e-9])
def test_contact_at_the_touching_distance(harness, rel):
    """Segment distances at (R_A + R_B)(1 +- rel)."""
    rng = np.random.default_rng(71)
    a, b = dome_pairs(rng, 3000, reach=3.0)
    target = 2 * RADIUS * (1.0 + rel * rng.choice([-1.0, 1.0], len(a)))
    a, b = at_distance(a, b, target)
    check_apart(harness, a, b)


def test_contact_near_parallel_and_crossing(harness):
    rng = np.random.default_rng(81)
    n = 5000
    a, _ = dome_pairs(rng, n)
    n = len(a)
    ua = unit(a["q"] - a["p"])
    ang = 10.0 ** rng.uniform(-8, -2, n)
    ub = unit(ua + ang[:, None] * unit(np.cross(ua, rng.normal(size=(n, 3)))))
    side = unit(np.cross(ua, rng.normal(size=(n, 3))))
    dist = 2 * RADIUS * rng.uniform(0.9, 1.1, n)
    dist[: n // 10] = 0.0  # crossing
    ca = 0.5 * (a["p"] + a["q"])
    b = cells_of(ca + dist[:, None] * side + rng.uniform(-3, 3, n)[:, None] * ua, ub, rng.uniform(2, 4, n))
    check_apart(harness, a, b)


def test_list_prefilter_keeps_every_end_point_within_reach(harness):
    """The list build's FP32 end-point prefilter (PreCol, pair_class.h) is
    conservative: every candidate with an end point within L_A/2 + Rc R_A of
    A's centre (the FP64 first test of reduceNeighbours, ABM/Neighbours.cpp:
    69-73) passes it, here for end points placed just inside the reach, with
    centres anywhere near the candidate's bin column and z up to 60 um. It is
    also a filter within its stated band: a kept candidate has an end point
    with |e - c|^2 <= 1.0001 r'^2 + 4e-6 E^2 (r' = reach + 1e-3, E =
    |c'_x| + |c'_y| + |c'_z| + 2 r') up to the FP32 rounding (2e-6 E^2 here)."""
    rng = np.random.default_rng(11)
    n = 1_000_000
    c = np.column_stack([rng.uniform(-10.0, 15.0, n), rng.uniform(-10.0, 15.0, n), rng.uniform(0.0, 60.0, n)])
    re = rng.uniform(0.5, 4.0, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    # near end point: just inside the reach for the first half, anywhere within
    # 1.5 reach for the second; the far one a cell length further
    t = np.where(np.arange(n) < n // 2, re * (1.0 - rng.uniform(0.0, 1e-6, n)), re * rng.uniform(0.0, 1.5, n))
    p = c + u * t[:, None]
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    q = p + v * rng.uniform(0.5, 6.0, n)[:, None]
    ends = np.ascontiguousarray(np.hstack([p, q]))
    centre = np.ascontiguousarray(c)
    reach = np.ascontiguousarray(re)
    out = np.zeros(n, np.int32)
    dp = ctypes.POINTER(ctypes.c_double)
    harness.pc_prefilter_batch(n, ends.ctypes.data_as(dp), centre.ctypes.data_as(dp), reach.ctypes.data_as(dp),
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    d = np.minimum(np.linalg.norm(p - c, axis=1), np.linalg.norm(q - c, axis=1))
    inside = d <= re
    assert inside.sum() > n // 2
    assert out[inside].all(), np.flatnonzero(inside & (out == 0))[:5]
    rr = re + 1e-3
    E = np.abs(c).sum(axis=1) + 2 * rr
    far = d**2 > 1.0001 * rr**2 + 6e-6 * E**2
    assert far.sum() > n // 10
    assert not out[far].any(), np.flatnonzero(far & (out == 1))[:5]
