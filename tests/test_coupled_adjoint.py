"""The full coupled configuration at reduced size: FP64 mechanics of a
growing, dividing colony; every second step a nutrient update whose sources
come from the moving cells (bcg_rd3d_deposit_cells), solved by the mixed
FP32-in-FP64 solver to 1e-10, with the growth rates of the next steps taken
from the new field; then the loss of the last field backpropagated through
every update by the checkpointed rollout (recomputation between
checkpoints), in mixed precision.

Checked link by link against independent references:
  * cells: the C oracle runs the same mechanics steps from the same state with
    the growth-rate fields the GPU produced; lineage exact, positions 1e-10;
  * field: each update against the oracle's solve (oracle/rd3d.py, conjugate
    gradients to 1e-14, pinned there against the direct sparse solve) of the
    step with the oracle's own cloud-in-cell deposition of the oracle's cells;
  * gradients (D, K, every update's per-cell rates): against the oracle's
    step-by-step adjoint chain of transpose solves.
"""
import numpy as np
import pytest

from oracle import rd3d

pytestmark = pytest.mark.gpu

DT_H = 1.0 / 3600.0


def say(*a):
    print(*a, flush=True)


def _radial_target(shape, h, origin):
    nz, ny, nx = shape
    x = origin[0] + h * np.arange(nx)
    y = origin[1] + h * np.arange(ny)
    r = np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)
    return np.ascontiguousarray(np.broadcast_to(20.0 * (1.0 - 0.8 * np.exp(-r / 150.0)), shape))


def test_coupled_mixed_updates_and_checkpointed_adjoint(oracle):
    import torch

    from baccolonygrowth_amd.adjoint import gather, rollout_seq
    from baccolonygrowth_amd.colony import dome_colony, dome_params, synthetic_fields
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params
    from baccolonygrowth_amd.simulation import Colony

    n = 20000
    cells = dome_colony(n, radius=110.0, height=12.0, seed=5)
    p = dome_params(cells, bin_size=5.0, max_neighbours=96, extra_cells=n)
    rng = np.random.default_rng(5)
    mu, K0, D0 = 0.875, 0.1, 600.0
    # cells at spread points of their cycle: divisions in every few steps
    cells["Ldiv"] = cells["Length"] * np.exp(1.585 * mu * DT_H * rng.uniform(1.0, 400.0, n))
    fields = synthetic_fields(p, cells, seed=5)
    nz, ny, nx, h = 17, 65, 65, 4.0
    origin = (-128.0, -128.0, 0.0)
    steps, every = 10, 2
    dt_nut = every * DT_H * 3600.0
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=h, D=D0, dt=dt_nut, K=K0, c_res=20.0, origin=origin, max_levels=4)
    rate = 1.66 / 3.0
    shape = (nz, ny, nx)
    say(f"[gpu] coupled config: {n} cells, {steps} steps, mixed update every {every} on {nx}x{ny}x{nz}")
    o = oracle.sim(p)
    o.set_cells(cells)
    o.set_fields(*fields)
    oracle.srand(5)
    seq_cm, seq_rate, fields_run = [], [], []
    with Colony(p) as col, Nutrients3D(prm) as s:
        col.LoadCells(cells)
        col.set_fields(*fields)
        col.srand(5)
        c_prev = s.field()
        c_init = c_prev.copy()
        t, ndiv = 0.0, 0
        for k in range(steps):
            la = o.step(DT_H, t, rebuild=True)
            lb = col.step(DT_H, t, update=True, no_patch=True)
            assert np.array_equal(la, lb), f"step {k}: division events differ"
            ndiv += len(lb)
            t += DT_H
            if (k + 1) % every:
                continue
            s.deposit_cells(col, rate)
            it, rr = s.step("mixed", tol=1e-10, max_iter=100)
            assert rr <= 1e-10
            s.growth_to_cells(col, mu, K0)
            gr1 = col.fields()[3]
            o.set_fields(gr1=gr1)
            c_new = s.field()
            # this update's sources: the GPU cells, and the oracle's cells
            cb, ca = col.cells(), o.cells()
            cm_b = np.ascontiguousarray((cb["p"] + cb["q"]) * 0.5)
            cm_a = (ca["p"] + ca["q"]) * 0.5
            assert np.abs(cm_b - cm_a).max() <= 1e-10 * np.abs(cm_a).max()
            seq_cm.append(cm_b)
            seq_rate.append(rate * cb["Length"])
            q_ref, _ = rd3d.deposit_cic(cm_a, rate * ca["Length"], shape, h, origin)
            c_ref = rd3d.step_direct(c_prev, q_ref, K0, D0, dt_nut, h, 20.0, solver="cg")
            err = np.abs(c_new - c_ref).max()
            # ||A^-1|| <= 1 (A = I + dt sigma - dt D Lap_h, sigma >= 0): the
            # solve to ||b - A c|| <= 1e-10 ||b|| is within 1e-10 ||b|| of the
            # exact step
            bound = 1e-10 * np.linalg.norm(rd3d.rhs(c_prev, 20.0)) + 1e-11
            say(f"    update {len(seq_cm)}: {len(cb)} cells, {it} refinements, relres {rr:.2e}, "
                f"|c - c_direct| {err:.2e} (bound {bound:.2e})")
            assert err <= bound
            assert c_new.min() < 20.0 - 1e-3  # the colony draws the field down
            fields_run.append(c_new)
            c_prev = c_new
        assert ndiv > 0
        ref_cells = o.cells()
        got = col.cells()
        assert len(got) == len(ref_cells) == n + ndiv
        for f in ("Type", "Ancestor", "age_p", "age_q"):
            assert np.array_equal(got[f], ref_cells[f]), f
        for f in ("p", "q"):
            assert np.abs(got[f] - ref_cells[f]).max() <= 1e-10 * np.abs(ref_cells[f]).max(), f
        assert col.rand() == oracle.rand()

        # the loss of the last field through every update, checkpoints every 2
        target = _radial_target(shape, h, origin)
        rates = [torch.tensor(r, requires_grad=True) for r in seq_rate]
        Dt = torch.tensor(D0, dtype=torch.float64, requires_grad=True)
        Kt = torch.tensor(K0, dtype=torch.float64, requires_grad=True)
        c_fin = rollout_seq(rates, Dt, Kt, torch.from_numpy(c_init), s, seq_cm, checkpoint_every=2, mode="mixed",
                            tol=1e-10, adj_tol=1e-12)
        # the rollout recomputes the run's fields exactly (same solves)
        assert c_fin.detach().numpy().tobytes() == fields_run[-1].tobytes()
        ((c_fin - torch.from_numpy(target)) ** 2).sum().backward()

        # oracle chain: forward fields of the run, direct transpose solves back
        qs = [rd3d.deposit_cic(seq_cm[k], seq_rate[k], shape, h, origin)[0] for k in range(len(seq_cm))]
        cs = [c_init] + fields_run
        g = 2.0 * (cs[-1] - target)
        dK_ref = dD_ref = 0.0
        dqs = [None] * len(qs)
        for k in range(len(qs) - 1, -1, -1):
            dq, dc0, dK, dD = rd3d.adjoint_direct(cs[k], cs[k + 1], qs[k], K0, D0, dt_nut, h, g, solver="cg")
            dqs[k] = dq
            dK_ref += dK
            dD_ref += dD
            g = dc0
        refs = [gather(s, seq_cm[k], dqs[k]) for k in range(len(qs))]
    say(f"    dL/dD {Dt.grad.item():.10e} (oracle {dD_ref:.10e}), dL/dK {Kt.grad.item():.10e} (oracle {dK_ref:.10e})")
    assert Dt.grad.item() == pytest.approx(dD_ref, rel=1e-7)
    assert Kt.grad.item() == pytest.approx(dK_ref, rel=1e-7)
    for k in range(len(qs)):
        e = np.abs(rates[k].grad.numpy() - refs[k]).max() / np.abs(refs[k]).max()
        assert e <= 1e-7, (k, e)
    o.close()
