"""Pins the oracle (oracle/bcg_oracle.c) before anything is checked against it.

* against the reference's golden vectors (tests/golden/*.npz, written by the
  reference's own code compiled from source) - bit-exact;
* against oracle/_ref (the reference library itself) where it is built;
* against mathematics: segment distance by brute force, closed forms of the
  Hertz contact, the glibc rand() stream of the host C library, invariants
  of the 3-D generalisation.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from baccolonygrowth_amd.cells import CELL_DTYPE
from helpers import GOLDEN, demo_params, planar_case


def _cells(a):
    return np.ascontiguousarray(a).view(CELL_DTYPE).reshape(-1)


def _eq(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


# ---------------------------------------------------------------- RNG ---
def test_glibc_stream_matches_host_libc(oracle):
    libc = ctypes.CDLL(None)
    for seed in (1, 2, 12345, 2**31 - 2):
        libc.srand(seed)
        host = np.array([libc.rand() for _ in range(1000)])
        assert _eq(oracle.glibc_stream(seed, 1000), host)


def test_glibc_known_answers(oracle):
    # published first outputs of glibc rand() after srand(1)
    assert oracle.glibc_stream(1, 3).tolist() == [1804289383, 846930886, 1681692777]


# ------------------------------------------------------- pair functions ---
@pytest.mark.parametrize("tag", ["planar", "3d"])
def test_min_distance_and_F_cc_golden(oracle, tag):
    z = np.load(os.path.join(GOLDEN, "golden_pairs.npz"))
    a, b = _cells(z[f"{tag}_a"]), _cells(z[f"{tag}_b"])
    p = demo_params()
    md = np.stack([oracle.min_distance(a[i : i + 1], b[i : i + 1]) for i in range(len(a))])
    fcc = np.stack([oracle.F_cc(p, a[i : i + 1], b[i : i + 1]) for i in range(len(a))])
    assert _eq(md, z[f"{tag}_min_distance"])
    assert _eq(fcc, z[f"{tag}_F_cc"])


def test_F_cw_golden(oracle):
    z = np.load(os.path.join(GOLDEN, "golden_pairs.npz"))
    c = _cells(z["cw_cells"])
    p = demo_params()
    oracle.srand(5)
    out = np.stack([oracle.F_cw(p, c[i : i + 1], z["cw_wall_z"][i]) for i in range(len(c))])
    assert _eq(out, z["cw_out"])


def _seg_points(c, n=401):
    s = np.linspace(0, 1, n)[:, None]
    return c["p"][0] + s * (c["q"][0] - c["p"][0])


def test_min_distance_brute_force(oracle):
    """d equals the minimum over dense samples of both segments (to the
    sampling resolution) and is never larger than any sampled distance."""
    z = np.load(os.path.join(GOLDEN, "golden_pairs.npz"))
    a, b = _cells(z["3d_a"]), _cells(z["3d_b"])
    for i in range(0, 400, 7):
        out = oracle.min_distance(a[i : i + 1], b[i : i + 1])
        pa, pb = _seg_points(a[i : i + 1]), _seg_points(b[i : i + 1])
        dd = np.sqrt(((pa[:, None, :] - pb[None, :, :]) ** 2).sum(-1))
        step = max(a["Length"][i], b["Length"][i]) / 400
        assert out[0] <= dd.min() + 1e-12
        assert out[0] >= dd.min() - step
        # the returned closest points realise the distance
        assert abs(np.linalg.norm(out[1:4] - out[4:7]) - out[0]) < 1e-12


def test_F_cc_closed_form(oracle):
    """Head-on overlap with no relative motion: |F| = sqrt(R1R2/(R1+R2))
    k_cc delta^1.5 along the line of centres; no force when apart."""
    p = demo_params()
    a = np.zeros(1, CELL_DTYPE)
    b = np.zeros(1, CELL_DTYPE)
    a["p"], a["q"], a["Length"], a["Radius"] = (1, 0, 1), (-1, 0, 1), 2.0, 0.5
    for gap in (0.9, 0.7, 1.2):
        b["p"], b["q"], b["Length"], b["Radius"] = (1, 0, 1 + gap), (-1, 0, 1 + gap), 2.0, 0.5
        out = oracle.F_cc(p, a, b)
        delta = max(1.0 - gap, 0.0)
        fe = math.sqrt(0.25 / 1.0) * p.k_cc * delta**1.5
        assert out[6] == pytest.approx(gap, rel=0, abs=1e-15)
        assert out[2] == pytest.approx(-fe, rel=1e-14, abs=0)
        assert out[0] == 0.0 and out[1] == 0.0
        # contact point halfway between the axes for equal radii
        assert out[5] == pytest.approx(1 + gap / 2, abs=1e-14)


def test_F_cc_pair_symmetry(oracle):
    """Newton's third law for the elastic part (no velocities)."""
    z = np.load(os.path.join(GOLDEN, "golden_pairs.npz"))
    a, b = _cells(z["3d_a"]).copy(), _cells(z["3d_b"]).copy()
    for c in (a, b):
        c["Velocity"] = 0
        c["AngularVelocity"] = 0
    p = demo_params()
    for i in range(0, 400, 3):
        f1 = oracle.F_cc(p, a[i : i + 1], b[i : i + 1])
        f2 = oracle.F_cc(p, b[i : i + 1], a[i : i + 1])
        if f1[6] == 0.0:  # intersecting axes: the normal is 0/0 (NaN), as in the reference
            assert np.isnan(f1[:3]).all()
            continue
        scale = max(np.abs(f1[:3]).max(), 1e-30)
        assert np.abs(f1[:3] + f2[:3]).max() <= 1e-12 * scale


# -------------------------------------------------- whole-step parity ---
@pytest.mark.parametrize("k", [7, 8])
def test_step_golden(oracle, k):
    z = np.load(os.path.join(GOLDEN, f"golden_step_r{k}.npz"))
    p = demo_params()
    sim = oracle.sim(p)
    sim.set_cells(_cells(z["cells_in"]))
    sim.set_fields(z["height"], z["normal"], z["wall"], z["gr1"], z["gr2"])
    sim.rebuild()
    c0, l0 = sim.neighbours(stride=p.maxNeighbours)
    assert _eq(c0, z["nl_counts0"])
    for i in range(len(c0)):
        assert _eq(l0[i, : c0[i]], z["nl_lists0"][i, : c0[i]])
    F, T = sim.sum_forces()
    assert _eq(F, z["F0"]) and _eq(T, z["T0"])
    oracle.srand(int(z["seed"]))
    t = float(z["t"])
    lin = []
    for s in range(int(z["steps"])):
        L = sim.step(float(z["dt"]), t, rebuild=(s == int(z["rebuild_step"])))
        lin.append(np.column_stack([np.full(len(L), s), L]))
        t += float(z["dt"])
    assert _eq(np.concatenate(lin), z["lineage"])
    assert _eq(sim.cells().view(np.uint8).reshape(-1, 208), z["cells_out"])
    c1, l1 = sim.neighbours(stride=p.maxNeighbours)
    assert _eq(c1, z["nl_counts1"])
    for i in range(len(c1)):
        assert _eq(l1[i, : c1[i]], z["nl_lists1"][i, : c1[i]])
    assert [oracle.rand() for _ in range(4)] == z["next_draws"].tolist()
    sim.close()


def test_oracle_vs_reference_library(oracle, reflib):
    """Same state through the restatement and the reference code: boxes,
    lists, forces and 6 steps (with two list rebuilds) agree bit for bit."""
    p, t, dt, cells, fields = planar_case(5, seed=55, divide_frac=0.2)
    sims = [oracle.sim(p), reflib.sim(p)]
    for s in sims:
        s.set_cells(cells)
        s.set_fields(*fields)
        s.rebuild()
    b0, b1 = sims[0].boxes(), sims[1].boxes()
    assert _eq(b0[0], b1[0])
    out = []
    for s, lib in zip(sims, (oracle, reflib)):
        lib.srand(99)
        tt = t
        lins = []
        for k in range(6):
            lins.append(s.step(dt, tt, rebuild=(k % 3 == 0)))
            tt += dt
        out.append((s.cells(), s.neighbours(), s.sum_forces(), np.concatenate(lins), lib.rand()))
    (ca, na, fa, la, ra), (cb, nb, fb, lb, rb) = out
    assert _eq(la, lb) and len(la) > 0
    assert ca.tobytes() == cb.tobytes()
    assert _eq(na[0], nb[0]) and _eq(na[1], nb[1])
    assert _eq(fa[0], fb[0]) and _eq(fa[1], fb[1])
    assert ra == rb


def test_mean_stress_oracle_vs_reference(oracle, reflib):
    """mean_stress of every cell, as the output step computes it: the
    restatement and the reference code agree bit for bit, and take the same
    rand() draws (one per F_cw call)."""
    p, t, dt, cells, fields = planar_case(5, seed=56, divide_frac=0.2)
    out = []
    for lib in (oracle, reflib):
        s = lib.sim(p)
        s.set_cells(cells)
        s.set_fields(*fields)
        s.rebuild()
        lib.srand(7)
        S, F = s.mean_stress()
        out.append((S, F, lib.rand()))
        s.close()
    (Sa, Fa, ra), (Sb, Fb, rb) = out
    assert Sa.tobytes() == Sb.tobytes() and Fa.tobytes() == Fb.tobytes()
    assert ra == rb
    assert np.abs(Sa[:, :3]).max() > 0 and np.abs(Fa).max() > 0
    oracle.srand(7)
    assert ra != oracle.rand()  # the F_cw calls took draws


def test_cells_text_output_matches_reference(oracle, reflib, tmp_path):
    """The output step's cell file written from the oracle's cells and
    mean_stress equals the one the reference's own mean_stress and Output
    write for the same state, byte for byte."""
    from baccolonygrowth_amd.cells import write_cells_text

    p, t, dt, cells, fields = planar_case(5, seed=57, divide_frac=0.2)
    sims = [oracle.sim(p), reflib.sim(p)]
    for s, lib in zip(sims, (oracle, reflib)):
        s.set_cells(cells)
        s.set_fields(*fields)
        s.rebuild()
        lib.srand(11)  # one libc stream: seeded before each library's step
        s.step(dt, t, rebuild=False)  # velocities and frictions of a step
    ref_path, ours_path = tmp_path / "ref.txt", tmp_path / "ours.txt"
    sims[1].output_cells(str(ref_path), t + dt)
    S, _ = sims[0].mean_stress()
    write_cells_text(str(ours_path), t + dt, sims[0].cells(), S)
    ref, ours = ref_path.read_bytes(), ours_path.read_bytes()
    assert ref.count(b"\n") == sims[1].n >= len(cells) and len(ref.split(b"\n")[0].split()) == 28
    assert ours == ref
    for s in sims:
        s.close()


# ------------------------------------------- 3-D generalisation pinning ---
def _embed_3d(p, cells, fields, ny=5):
    """A planar state inside a 3-D grid (BoxY = ny) with y-constant fields."""
    q = p.replace(planar=0, BoxY=ny)
    h, n, w, g1, g2 = fields
    ref = p.refinementGridHeight
    h3 = np.repeat(h, ny * ref, axis=1)
    n3 = np.repeat(n, ny * ref, axis=1)
    w3 = np.repeat(w, ny, axis=1)
    g13 = np.repeat(g1, ny, axis=1)
    g23 = np.repeat(g2, ny, axis=1)
    return q, (h3, n3, w3, g13, g23)


def test_3d_mode_reduces_to_planar(oracle):
    """With every cell, velocity and contact in the x-z plane and no
    out-of-plane spin, the 3-D model equals the planar model: identical
    neighbour sets, positions within 1e-12 (the y weights of the bilinear
    height lookup round differently). No divisions: a daughter is born
    with a random spin about z (grow.cpp:152), which leaves the plane."""
    p, t, dt, cells, fields = planar_case(7, seed=3, divide_frac=0.0)
    cells["AngularVelocity"][:, 0] = 0
    cells["AngularVelocity"][:, 2] = 0
    p = p.replace(varAngle=0.0)
    # keep the y bin index central: with BoxY = 5 the cell row y=0 maps to
    # int(0 + 2 + 0.5) = 2, so the 27-box gather sees the same boxes
    q, f3 = _embed_3d(p, cells, fields)
    res = []
    for prm, flds in ((p, fields), (q, f3)):
        s = oracle.sim(prm)
        s.set_cells(cells)
        s.set_fields(*flds)
        s.rebuild()
        oracle.srand(7)
        lins = [s.step(dt, t + k * dt, rebuild=(k == 1)) for k in range(3)]
        res.append((s.cells(), s.neighbours(), np.concatenate(lins)))
        s.close()
    (ca, na, la), (cb, nb, lb) = res
    assert _eq(la, lb)
    assert _eq(na[0], nb[0])
    for i in range(len(na[0])):
        assert sorted(na[1][i, : na[0][i]]) == sorted(nb[1][i, : nb[0][i]])
    assert np.abs(cb["p"][:, 1]).max() < 1e-12
    scale = np.abs(ca["p"]).max()
    assert np.abs(ca["p"] - cb["p"]).max() <= 1e-12 * scale
    assert np.abs(ca["Velocity"] - cb["Velocity"]).max() <= 1e-9 * max(np.abs(ca["Velocity"]).max(), 1)


def test_3d_rotation_about_z_invariance(oracle):
    """Rotating a 3-D state by 90 degrees about z (x->y, y->-x) with
    rotation-symmetric fields gives the rotated forces."""
    p, t, dt, cells, fields = planar_case(5, seed=5, divide_frac=0.0)
    q, f3 = _embed_3d(p, cells, fields, ny=p.BoxX)
    q = q.replace(tension=0.0, wall_rough=0.0)
    h, n, w, g1, g2 = f3
    w[:] = 0.0
    rc = cells.copy()
    for key in ("p", "q", "Velocity", "AngularVelocity"):
        v = cells[key].copy()
        rc[key][:, 0] = -v[:, 1]
        rc[key][:, 1] = v[:, 0]
    F = []
    for cc in (cells, rc):
        s = oracle.sim(q)
        s.set_cells(cc)
        s.set_fields(h, n, w, g1, g2)
        s.rebuild()
        F.append(s.sum_forces())
        s.close()
    (Fa, Ta), (Fb, Tb) = F
    scale = np.abs(Fa).max()
    assert np.abs(Fb[:, 0] + Fa[:, 1]).max() <= 1e-9 * scale
    assert np.abs(Fb[:, 1] - Fa[:, 0]).max() <= 1e-9 * scale
    assert np.abs(Fb[:, 2] - Fa[:, 2]).max() <= 1e-9 * scale


@pytest.mark.parametrize("case", ["planar", "dome"])
def test_openmp_step_equals_serial_step(oracle, case):
    """The OpenMP form of the oracle step (the timed CPU baseline) gives the
    serial step's cells, lists, lineage and rand() position bit for bit."""
    from baccolonygrowth_amd.colony import dome_colony, dome_params, synthetic_fields

    if case == "planar":
        p, t, dt, cells, fields = planar_case(5, seed=58, divide_frac=0.2)
    else:
        cells = dome_colony(6000, radius=60.0, height=10.0, seed=9)
        p = dome_params(cells, bin_size=5.0, max_neighbours=96, extra_cells=6000)
        rng = np.random.default_rng(9)
        cells["Ldiv"] = cells["Length"] * (1.0 + rng.uniform(0, 3e-3, len(cells)))
        fields = synthetic_fields(p, cells, seed=9)
        t, dt = 0.0, 1.0 / 3600
    out = []
    for par in (False, True):
        s = oracle.sim(p)
        s.set_cells(cells)
        s.set_fields(*fields)
        oracle.srand(21)
        oracle.set_threads(4)
        lins, tt = [], t
        for k in range(5):
            step = s.step_par if par else s.step
            lins.append(step(dt, tt, rebuild=(k % 2 == 0)))
            tt += dt
        out.append((s.cells(), s.neighbours(), np.concatenate(lins), oracle.rand()))
        s.close()
    (ca, na, la, ra), (cb, nb, lb, rb) = out
    assert len(la) > 0 and _eq(la, lb)
    assert ca.tobytes() == cb.tobytes()
    assert _eq(na[0], nb[0]) and _eq(na[1], nb[1])
    assert ra == rb
