"""3-D nutrient step: the FP64 oracle (oracle/rd3d.py) pinned by mathematics
on the CPU, then the HIP multigrid (FP64 and mixed precision) and the
deposition checked against it on the GPU.

Tolerances: the solve stops at ||b - A c|| <= 1e-10 ||b||; the
solution is compared with a direct sparse solve at 1e-9 relative (the
residual tolerance times the operator's conditioning on these grids)."""
import numpy as np
import pytest

from oracle import rd3d


def _case(shape, seed=0, lam=50.0, dt=10.0, K=0.1):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    c_old = 20.0 * (1.0 - 0.5 * rng.random(shape))
    c_old[0] = 20.0
    q = np.zeros(shape)
    q[1 : nz // 2, ny // 4 : 3 * ny // 4, nx // 4 : 3 * nx // 4] = rng.uniform(0.0, 0.05, (nz // 2 - 1, 3 * ny // 4 - ny // 4, 3 * nx // 4 - nx // 4))
    sigma = q / (K + c_old)
    return c_old, q, sigma


# ----------------------------------------------------- oracle pinning (CPU)
def test_operator_two_constructions_agree():
    shape = (5, 6, 7)
    c_old, q, sigma = _case(shape)
    A = rd3d.sparse_A(shape, sigma, 3.0, 10.0)
    u = np.random.default_rng(1).random(shape)
    assert np.allclose(A @ u.ravel(), rd3d.apply_A(u, sigma, 3.0, 10.0).ravel(), rtol=1e-14, atol=1e-12)


def test_operator_symmetrised_and_constants():
    shape = (5, 6, 7)
    _, _, sigma = _case(shape)
    A = rd3d.sparse_A(shape, sigma, 3.0, 10.0).toarray()
    W = rd3d.control_volume_weights(shape).ravel()
    free = np.arange(A.shape[0]) >= shape[1] * shape[2]  # not Dirichlet
    S = (W[:, None] * A)[np.ix_(free, free)]
    assert np.abs(S - S.T).max() < 1e-12
    # constants are in the kernel of the diffusion part
    one = np.ones(shape)
    assert np.abs(rd3d.apply_A(one, 0 * sigma, 3.0, 10.0) - 1.0).max() < 1e-13


def test_direct_solve_and_manufactured():
    shape = (9, 9, 9)
    c_old, q, sigma = _case(shape)
    b = rd3d.rhs(c_old, 20.0)
    x = rd3d.solve_direct(b, sigma, 40.0, 10.0)
    assert np.abs(rd3d.residual(b, x, sigma, 40.0, 10.0)).max() < 1e-10
    ustar = np.random.default_rng(3).random(shape)
    bs = rd3d.apply_A(ustar, sigma, 40.0, 10.0)
    assert np.abs(rd3d.solve_direct(bs, sigma, 40.0, 10.0) - ustar).max() < 1e-10


def test_deposit_cic_conserves_and_localises():
    rng = np.random.default_rng(0)
    shape = (6, 7, 8)
    cm = rng.uniform([0, 0, 0], [6.9, 5.9, 4.9], (500, 3))
    rate = rng.random(500)
    q, inside = rd3d.deposit_cic(cm, rate, shape, 1.0, (0, 0, 0))
    assert inside.all()
    assert abs(q.sum() - rate.sum()) < 1e-12
    q1, _ = rd3d.deposit_cic(np.array([[2.0, 3.0, 1.0]]), np.array([2.0]), shape, 1.0, (0, 0, 0))
    assert q1[1, 3, 2] == 2.0 and q1.sum() == 2.0


# ------------------------------------------------------------- GPU tests
def _solver(shape, h=1.0, D=600.0, dt=10.0, K=0.1, **kw):
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params

    nz, ny, nx = shape
    p = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=h, D=D, dt=dt, K=K, c_res=20.0, origin=(0.0, 0.0, 0.0), **kw)
    return Nutrients3D(p)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(17, 33, 33), (9, 17, 33), (5, 9, 9)])
def test_step_matches_direct_solve(mode, shape):
    print(f"[gpu] rd3d {mode} {shape}", flush=True)
    h, D, dt, K = 1.0, 600.0, 10.0, 0.1
    c_old, q, sigma = _case(shape, seed=4, K=K)
    lam = dt * D / h**2
    b = rd3d.rhs(c_old, 20.0)
    ref = rd3d.solve_direct(b, sigma, lam, dt)
    r0 = np.linalg.norm(b)
    with _solver(shape, h, D, dt, K) as s:
        s.set_field(c_old)
        s.set_uptake(q)
        it, rr = s.step(mode, tol=1e-10, max_iter=60)
        c = s.field()
    res = np.linalg.norm(rd3d.residual(b, c, sigma, lam, dt)) / r0
    err = np.linalg.norm(c - ref) / np.linalg.norm(ref)
    print(f"    iters {it}, reported {rr:.2e}, numpy residual {res:.2e}, error vs direct {err:.2e}", flush=True)
    assert res <= 1.05e-10
    assert abs(res - rr) <= 1e-3 * rr + 1e-14
    assert err <= 1e-9
    assert it < 40


@pytest.mark.gpu
def test_mixed_equals_fp64_and_deterministic():
    shape = (17, 33, 33)
    c_old, q, _ = _case(shape, seed=7)
    outs = {}
    for mode in ("fp64", "mixed", "mixed"):
        with _solver(shape) as s:
            s.set_field(c_old)
            s.set_uptake(q)
            s.step(mode, tol=1e-10)
            c = s.field()
        if mode in outs:
            assert c.tobytes() == outs[mode].tobytes()  # bitwise repeatable
        outs[mode] = c
    d = np.abs(outs["mixed"] - outs["fp64"]).max() / np.abs(outs["fp64"]).max()
    print(f"[gpu] mixed vs fp64 max rel difference {d:.2e}", flush=True)
    assert d <= 1e-9


@pytest.mark.gpu
def test_deposit_matches_oracle_and_is_deterministic():
    shape = (9, 17, 33)
    rng = np.random.default_rng(5)
    n = 20000
    cm = rng.uniform([-2.0, -2.0, -1.0], [34.0, 18.0, 10.0], (n, 3))
    rate = rng.random(n)
    ref, inside = rd3d.deposit_cic(cm, rate, shape, 1.0, (0.0, 0.0, 0.0))
    with _solver(shape) as s:
        ni = s.deposit(cm, rate)
        q1 = s.uptake()
        s.deposit(cm, rate)
        q2 = s.uptake()
    assert ni == int(inside.sum())
    assert q1.tobytes() == q2.tobytes()
    assert np.abs(q1 - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.gpu
def test_full_size_257_converges_both_modes():
    """257x257x65 at 2 um (the kernel-bench grid): both solvers reach
    1e-10, checked by an independent numpy residual, and agree."""
    shape = (65, 257, 257)
    print("[gpu] rd3d 257x257x65", flush=True)
    rng = np.random.default_rng(9)
    c_old = np.full(shape, 20.0)
    q = np.zeros(shape)
    q[1:12, 64:192, 64:192] = rng.uniform(0, 0.05, (11, 128, 128))
    h, D, dt, K = 2.0, 600.0, 10.0, 0.1
    sigma = q / (K + c_old)
    lam = dt * D / h**2
    b = rd3d.rhs(c_old, 20.0)
    r0 = np.linalg.norm(b)
    out = {}
    for mode in ("fp64", "mixed"):
        with _solver(shape, h, D, dt, K) as s:
            s.set_field(c_old)
            s.set_uptake(q)
            it, rr = s.step(mode, tol=1e-10, max_iter=60)
            out[mode] = s.field()
        res = np.linalg.norm(rd3d.residual(b, out[mode], sigma, lam, dt)) / r0
        print(f"    {mode}: {it} iterations, residual {res:.2e}", flush=True)
        assert res <= 1.05e-10
    d = np.abs(out["mixed"] - out["fp64"]).max() / 20.0
    assert d <= 1e-9
