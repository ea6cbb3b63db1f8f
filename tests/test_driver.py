"""The drop-in driver (baccolonygrowth_amd.driver): the reference
executable's sequence - parameters, AddFirstCells or a restart (cells and
fields), the rough wall, the initial pass, the time loop with its output
tree and lineage file.

CPU: the pieces the driver adds, against the reference compiled from source
(oracle/_ref): the first cell and its rand() draw, the rough wall's float
arithmetic and draws. GPU: a restart run through the driver against the
reference's own run of the same sequence (tests/golden/golden_driver.npz,
written by tests/golden/make_env_golden.py driver): lineage exactly; saved
cells, final cells at 1e-10 as the mechanics tests; the text files
number by number to their printed precision.
"""
import lzma
import os

import numpy as np
import pytest

from baccolonygrowth_amd.cells import CELL_DTYPE, first_cells, load_restart, read_parameters
from helpers import GOLDEN, demo_params, driver_input

REF_SO = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "libbcgref.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref (the reference compiled from source) "
                               "is not built here")
GOLDEN_DRIVER = os.path.join(GOLDEN, "golden_driver.npz")


@needs_ref
def test_first_cell_and_its_draw_match_add_first_cells():
    from oracle.oracle import Oracle, RefLib

    p = demo_params()
    p.maxCells = 16
    ref, orc = RefLib(), Oracle()
    sim = ref.sim(p)
    ref.srand(5)
    assert ref.add_first_cells(sim, 1, 2.0) == 1
    orc.srand(5)
    ours = first_cells(p, 1, rand=orc.rand)
    assert ours.tobytes() == sim.cells().tobytes()
    assert ref.rand() == orc.rand()  # one draw taken on each side
    sim.close()


def test_first_cells_rejects_a_colony_the_reference_cannot_place():
    with pytest.raises(ValueError):
        first_cells(demo_params(), 16)


@needs_ref
@pytest.mark.parametrize("rough", [0.1, 0.37])
def test_rough_wall_matches_reference_expression(rough):
    """((float)rand()/RAND_MAX - 0.5) * wall_rough over the same stream: the
    float quotient (not double) and one draw per column."""
    from baccolonygrowth_amd.driver import rough_wall
    from oracle.oracle import Oracle, RefLib

    p = demo_params()
    p.wall_rough = rough
    ref, orc = RefLib(), Oracle()
    ref.srand(123)
    exp = ref.rough_wall(p)
    orc.srand(123)
    got = rough_wall(p, orc.rand)
    assert got.tobytes() == exp.tobytes()
    assert ref.rand() == orc.rand()
    # the double quotient would differ
    orc.srand(123)
    dbl = np.array([(orc.rand() / 2147483647.0 - 0.5) * rough for _ in range(p.BoxX)])
    assert not np.array_equal(dbl, exp)


def test_driver_input_keys(tmp_path):
    inp = tmp_path / "in.txt"
    driver_input(str(inp), "/some/dir", 6.01)
    p, extra = read_parameters(str(inp))
    assert (p.BoxX, p.BoxZ, p.maxCells) == (128, 16, 4000)
    assert extra["Restart_dir"] == "/some/dir" and int(extra["Restart"]) == 1 and int(extra["RestartIndex"]) == 5
    assert extra["t_max"] == 6.01 and extra["Output_Time"] == 0.001


def _numbers(raw: bytes):
    lines = raw.decode().split("\n")
    return [np.array([float(v) for v in ln.split()]) for ln in lines]


def _same_text(got: bytes, exp: bytes, rtol: float, atol: float, label: str, skip=None):
    """Same lines and tokens; numbers equal to the printed precision (a
    last-digit difference allowed where the values differ at round-off).
    skip(line) gives a slice of a line's values not compared."""
    if got == exp:
        return 1.0
    a, b = _numbers(got), _numbers(exp)
    assert len(a) == len(b), label
    same = 0
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.shape == y.shape, (label, i)
        keep = np.ones(len(x), bool)
        if skip is not None and skip(i) is not None:
            keep[skip(i)] = False
        assert np.allclose(x[keep], y[keep], rtol=rtol, atol=atol), (label, i, np.abs(x - y)[keep].max())
        same += np.array_equal(x, y)
    return same / len(a)


def _wall_skip(bx: int):
    """WallField rows of levels >= 1 (four rows a level when BoxY = 1): the
    x range that the finer level covers holds a term the reference leaves
    uninitialised (its value is overwritten in every outer iteration)."""
    return lambda row: slice(bx // 4, bx * 3 // 4) if row >= 4 else None


@pytest.mark.gpu
def test_driver_restart_run_matches_reference(tmp_path):
    from baccolonygrowth_amd.driver import run
    from baccolonygrowth_amd.output import OUTPUT_DIRS

    g = np.load(GOLDEN_DRIVER)
    rs = tmp_path / "rs"
    for d in ("Restart", "Environment", "AgarField", "WallField"):
        (rs / d).mkdir(parents=True)
        (rs / d / "5.txt").write_bytes(lzma.decompress(g[f"rs_{d}"].tobytes()))
    inp = tmp_path / "input.txt"
    driver_input(str(inp), str(rs), float(g["t_max"]))
    out = tmp_path / "out"
    print(f"[gpu] driver restart run, {int(g['steps_run'])} steps", flush=True)
    s = run(str(inp), str(out), seed=int(g["seed"]), log=None)
    n_out = len(g["out_steps"])
    assert s["steps"] == int(g["steps_run"]) and s["outputs"] == list(range(n_out))
    for name in OUTPUT_DIRS:
        assert sorted(os.listdir(out / name)) == [f"{k}.txt" for k in range(n_out)], name
    exp_lineage = "".join("%d %d\n" % (a, b) for a, b in g["lineage"])
    assert (out / "lineage" / "0.txt").read_text() == exp_lineage
    for k in range(n_out):
        t, dt, moved = load_restart(str(out / "Restart" / f"{k}.txt"))
        exp = np.ascontiguousarray(g[f"out{k}_moved"]).view(CELL_DTYPE).reshape(-1)
        assert t == float(g[f"out{k}_t"]) and len(moved) == len(exp)
        for f in ("p", "q", "Length"):
            d = np.abs(moved[f] - exp[f]).max()
            assert d <= 1e-10 * max(np.abs(exp[f]).max(), 1.0), (k, f, d)
        frac = _same_text((out / "Cells" / f"{k}.txt").read_bytes(),
                          lzma.decompress(g[f"out{k}_cells_text"].tobytes()), 1e-4, 1.01e-4, f"Cells/{k}")
        assert frac >= 0.98, frac
    last = n_out - 1
    bx = read_parameters(str(inp))[0].BoxX
    for name in ("Density", "Environment", "AgarField", "WallField"):
        _same_text((out / name / f"{last}.txt").read_bytes(), lzma.decompress(g[f"last_{name}"].tobytes()),
                   1.5e-6, 1e-300, f"{name}/{last}", _wall_skip(bx) if name == "WallField" else None)
    got = s["final_cells"]
    exp = np.ascontiguousarray(g["cells"]).view(CELL_DTYPE).reshape(-1)
    assert len(got) == len(exp)
    for f in ("p", "q"):
        d = np.abs(got[f] - exp[f]).max()
        assert d <= 1e-10 * max(np.abs(exp[f]).max(), 1.0), (f, d)
    print(f"    ok: {s['divisions']} divisions, outputs {s['outputs']}", flush=True)
