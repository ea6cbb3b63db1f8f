"""The founding-colony configuration end to end on the device: 16 rods in a
5 um disk, FP64 mechanics with a nutrient update every 10 steps on
65 x 65 x 17 (deposition from the cells, implicit step to 1e-10, growth
rates), grown to 64 cells. The state it reaches is then checked against the
oracle: mechanics steps replayed from it by the serial C oracle, the
deposition by the oracle's cloud-in-cell, and the last solve's residual by
the oracle's sparse operator."""
import numpy as np
import pytest

from oracle import rd3d

pytestmark = pytest.mark.gpu

POS_RTOL = 1e-10


def test_founding_colony_grows_and_replays_on_the_oracle(oracle):
    from baccolonygrowth_amd.colony import growing_params, small_colony
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params
    from baccolonygrowth_amd.simulation import Colony

    cells = small_colony(16, seed=1)
    p = growing_params(128.0, 24.0, 2048)
    shape = (17, 65, 65)
    h, K, mu, rate, dt_h = 2.0, 0.1, 0.875, 0.55, 1.0 / 3600
    origin = (-64.0, -64.0, 0.0)
    prm = default_rd3d_params(nx=65, ny=65, nz=17, h=h, D=600.0, dt=10.0, K=K, c_res=20.0, origin=origin, max_levels=4)
    print("[gpu] founding colony to 64 cells", flush=True)
    with Colony(p) as col, Nutrients3D(prm) as s:
        col.LoadCells(cells)
        col.srand(1)
        s.growth_to_cells(col, mu, K)
        t, k = 0.0, 0
        while col.n < 64:
            assert k < 20000
            col.step(dt_h, t, update=True, no_patch=True)
            t += dt_h
            k += 1
            if k % 10 == 0:
                c_prev = s.field()
                s.deposit_cells(col, rate)
                q = s.uptake()
                cc = col.cells()
                s.step("fp64", tol=1e-10, max_iter=100)
                s.growth_to_cells(col, mu, K)
        c_new = s.field()
        state, fields = col.cells(), col.fields()
        print(f"    {col.n} cells after {k} steps", flush=True)
    # the last update: deposition and solve against the oracle
    cm = (cc["p"] + cc["q"]) * 0.5
    q_ref, _ = rd3d.deposit_cic(cm, rate * cc["Length"], shape, h, origin)
    assert np.abs(q - q_ref).max() <= 1e-13 * np.abs(q_ref).max()
    A = rd3d.sparse_A(shape, q / (K + c_prev), 600.0 * 10.0 / h**2, 10.0)
    b = rd3d.rhs(c_prev, 20.0)
    rel = np.linalg.norm(b.ravel() - A @ c_new.ravel()) / np.linalg.norm(b)
    assert rel <= 1e-10 and c_new.min() < 20.0
    # 20 mechanics steps from the state reached, GPU vs oracle
    o = oracle.sim(p)
    o.set_cells(state)
    o.set_fields(*fields)
    o.rebuild()
    with Colony(p) as col:
        col.LoadCells(state)
        col.set_fields(*fields)
        col.getNeighbours()
        oracle.srand(5)
        col.srand(5)
        nd = 0
        for i in range(20):
            la = o.step(dt_h, t, rebuild=True)
            lb = col.step(dt_h, t, update=True)
            assert np.array_equal(la, lb)
            nd += len(la)
            t += dt_h
        a, bb = col.cells(), o.cells()
        assert col.rand() == oracle.rand()
    o.close()
    assert len(a) == len(bb)
    for f in ("age_p", "age_q", "Type", "Ancestor"):
        assert np.array_equal(a[f], bb[f]), f
    for f in ("Length", "Ldiv", "GrowthRate"):
        assert np.abs(a[f] - bb[f]).max() <= 1e-12 * np.abs(bb[f]).max(), f
    for f in ("p", "q"):
        assert np.abs(a[f] - bb[f]).max() <= POS_RTOL * np.abs(bb[f]).max(), f
