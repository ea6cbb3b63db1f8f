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


def test_deposit_cic_first_moment():
    """Linear (cloud-in-cell) weights reproduce linear functions: for every
    axis, sum_v q_v x_v h^3 = sum_i rate_i x_i exactly (to rounding) for
    cells inside the grid, whatever their position in a cell."""
    rng = np.random.default_rng(1)
    shape, h, origin = (7, 9, 11), 2.5, (-3.0, 4.0, 1.0)
    nz, ny, nx = shape
    hi = np.array(origin) + h * (np.array([nx, ny, nz]) - 1)
    cm = rng.uniform(np.array(origin) - 2.0, hi + 2.0, (4000, 3))
    rate = rng.lognormal(0.0, 0.5, 4000)
    q, inside = rd3d.deposit_cic(cm, rate, shape, h, origin)
    assert 0 < inside.sum() < len(cm)
    z, y, x = np.meshgrid(*(origin[a] + h * np.arange(n) for a, n in ((2, nz), (1, ny), (0, nx))), indexing="ij")
    for axis, coord in ((0, x), (1, y), (2, z)):
        lhs = (q * coord).sum() * h**3
        rhs = (rate[inside] * cm[inside, axis]).sum()
        assert abs(lhs - rhs) <= 1e-11 * np.abs(rate[inside] * cm[inside, axis]).sum(), axis
    # a dropped index or a swapped weight breaks the moments: a cell at a
    # fractional position off the lattice moves the first moment by its offset
    q1, _ = rd3d.deposit_cic(np.array([[origin[0] + 1.3 * h, origin[1] + 2.6 * h, origin[2] + 0.2 * h]]),
                             np.array([1.0]), shape, h, origin)
    assert (q1 * x).sum() * h**3 == pytest.approx(origin[0] + 1.3 * h, rel=1e-13)
    assert (q1 * y).sum() * h**3 == pytest.approx(origin[1] + 2.6 * h, rel=1e-13)
    assert (q1 * z).sum() * h**3 == pytest.approx(origin[2] + 0.2 * h, rel=1e-13)


def test_cg_step_and_adjoint_equal_direct():
    """The conjugate-gradient forms (used on grids too large to factorise)
    against the direct sparse solves: the step and every adjoint output."""
    shape = (6, 7, 9)
    h, D, dt, K = 1.5, 30.0, 10.0, 0.1
    c_old, q, target = _loss_parts(shape, seed=5)
    a = rd3d.step_direct(c_old, q, K, D, dt, h, 20.0)
    b = rd3d.step_direct(c_old, q, K, D, dt, h, 20.0, solver="cg")
    assert np.abs(a - b).max() <= 1e-12 * 20
    g = 2 * (a - target)
    ra = rd3d.adjoint_direct(c_old, a, q, K, D, dt, h, g)
    rb = rd3d.adjoint_direct(c_old, a, q, K, D, dt, h, g, solver="cg")
    for u, v in zip(ra, rb):
        assert np.abs(np.asarray(u) - np.asarray(v)).max() <= 1e-11 * max(np.abs(np.asarray(u)).max(), 1e-300)


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
    assert abs(res - rr) <= 0.05 * rr + 1e-13  # both near the FP64 evaluation floor
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
    # ||A^-1|| <= 1 here (A = I + dt*sigma - dt*D*Lap_h with sigma >= 0), so two
    # iterates with residuals <= tol ||b|| differ by at most ~2 tol ||b||
    b = rd3d.rhs(c_old, 20.0)
    d = np.linalg.norm(outs["mixed"] - outs["fp64"]) / np.linalg.norm(b)
    print(f"[gpu] mixed vs fp64 relative difference {d:.2e}", flush=True)
    assert d <= 4e-10


@pytest.mark.gpu
@pytest.mark.parametrize("nu,shape", [((2, 2), (65, 161, 225)), ((1, 3), (65, 161, 225)), ((2, 2), (129, 353, 369))])
def test_fused_sweeps_equal_split_sweeps(nu, shape):
    """The z-marching two-colour sweep (the default on levels of >= 2^21
    points) gives the iterates of the one-launch-per-colour sweep, bitwise,
    in both solvers; odd sweep counts end with a split sweep. 225 x 161 x 65
    leaves partial 30 x 30 tiles on both x and y edges; 369 x 353 x 129
    (>= 2^24 points) runs its FP32 sweeps on 62 x 30 tiles, partial on both
    edges too, and with SMOOTH_FUSED_WIDE its FP64 sweeps as well."""
    from baccolonygrowth_amd.nutrients import SMOOTH_FUSED, SMOOTH_FUSED_WIDE, SMOOTH_SPLIT

    print(f"[gpu] rd3d fused vs split sweeps {shape} nu={nu}", flush=True)
    c_old, q, _ = _case(shape, seed=12)
    for mode in ("fp64", "mixed"):
        out = {}
        for sm in (SMOOTH_FUSED, SMOOTH_FUSED_WIDE, SMOOTH_SPLIT):
            with _solver(shape, nu1=nu[0], nu2=nu[1], smoother=sm) as s:
                s.set_field(c_old)
                s.set_uptake(q)
                it, rr = s.step(mode, tol=1e-10, max_iter=60)
                out[sm] = (s.field(), it, rr)
        (a, ia, ra), (b, ib, rb) = out[SMOOTH_FUSED], out[SMOOTH_SPLIT]
        print(f"    {mode}: fused {ia} it {ra:.3e}, split {ib} it {rb:.3e}, max |diff| {np.abs(a - b).max():.3e}", flush=True)
        assert ia == ib and a.tobytes() == b.tobytes()
        w, iw, _ = out[SMOOTH_FUSED_WIDE]
        assert iw == ib and w.tobytes() == b.tobytes()


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
    d = np.linalg.norm(out["mixed"] - out["fp64"]) / np.linalg.norm(b)
    print(f"    mixed vs fp64 relative difference {d:.2e}", flush=True)
    assert d <= 4e-10


# ----------------------------------------------------------- adjoint (CPU)
def _loss_parts(shape, seed=11):
    rng = np.random.default_rng(seed)
    c_old, q, _ = _case(shape, seed=seed)
    target = 20.0 * (1 - 0.3 * rng.random(shape))
    return c_old, q, target


def test_oracle_adjoint_matches_finite_differences():
    """adjoint_direct (sparse transpose solve) against central differences of
    the direct step: dL/dK, dL/dD, dL/dq and dL/dc_old at sample vertices."""
    shape = (5, 6, 7)
    h, D, dt, K = 1.0, 3.0, 10.0, 0.1
    c_old, q, target = _loss_parts(shape)

    def L(c_old=c_old, q=q, K=K, D=D):
        c = rd3d.step_direct(c_old, q, K, D, dt, h, 20.0)
        return float(((c - target) ** 2).sum())

    c1 = rd3d.step_direct(c_old, q, K, D, dt, h, 20.0)
    dq, dc0, dK, dD = rd3d.adjoint_direct(c_old, c1, q, K, D, dt, h, 2 * (c1 - target))
    e = 1e-3  # smaller steps are round-off limited on this tiny grid (truncation here ~e^2)
    assert dK == pytest.approx((L(K=K * (1 + e)) - L(K=K * (1 - e))) / (2 * e * K), rel=1e-4)
    assert dD == pytest.approx((L(D=D * (1 + e)) - L(D=D * (1 - e))) / (2 * e * D), rel=1e-4)
    for idx in [(1, 2, 3), (2, 3, 4), (3, 1, 2), (4, 5, 6)]:
        for arr, grad in ((q, dq), (c_old, dc0)):
            a1, a2 = arr.copy(), arr.copy()
            step = e * max(abs(arr[idx]), 1e-3)
            a1[idx] += step
            a2[idx] -= step
            kw1 = {"q": a1} if arr is q else {"c_old": a1}
            kw2 = {"q": a2} if arr is q else {"c_old": a2}
            fd = (L(**kw1) - L(**kw2)) / (2 * step)
            assert grad[idx] == pytest.approx(fd, rel=1e-5, abs=1e-9)


# ----------------------------------------------------------- adjoint (GPU)
@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_adjoint_matches_oracle(mode):
    from baccolonygrowth_amd import adjoint as adj

    shape = (9, 17, 17)
    h, D, dt, K = 1.0, 20.0, 10.0, 0.1
    c_old, q, target = _loss_parts(shape, seed=3)
    c1_ref = rd3d.step_direct(c_old, q, K, D, dt, h, 20.0)
    g = 2 * (c1_ref - target)
    ref = rd3d.adjoint_direct(c_old, c1_ref, q, K, D, dt, h, g)
    print(f"[gpu] adjoint {mode} {shape}", flush=True)
    with _solver(shape, h, D, dt, K) as s:
        s.set_field(c_old)
        s.set_uptake(q)
        s.step(mode, tol=1e-12, max_iter=200)
        c1 = s.field()
        dq, dc0, dK, dD = adj.adjoint(s, g, mode, tol=1e-12)
        c1_after = s.field()
    assert np.abs(c1 - c1_ref).max() <= 1e-10 * 20
    assert c1_after.tobytes() == c1.tobytes()  # forward state restored
    scale = np.abs(ref[0]).max()
    assert np.abs(dq - ref[0]).max() <= 1e-8 * scale
    assert np.abs(dc0 - ref[1]).max() <= 1e-8 * np.abs(ref[1]).max()
    assert dK == pytest.approx(ref[2], rel=1e-8)
    assert dD == pytest.approx(ref[3], rel=1e-8)


@pytest.mark.gpu
def test_failed_adjoint_keeps_the_forward_state():
    """An adjoint solve that cannot converge (one iteration to 1e-14) fails
    and leaves the forward state as it was: the field is still c_new and the
    next step equals the one a fresh solver makes from c_new."""
    from baccolonygrowth_amd import adjoint as adj
    from baccolonygrowth_amd._lib import BcgError

    shape = (9, 17, 17)
    h, D, dt, K = 1.0, 20.0, 10.0, 0.1
    c_old, q, target = _loss_parts(shape, seed=5)
    print(f"[gpu] failed adjoint {shape}", flush=True)
    with _solver(shape, h, D, dt, K) as a, _solver(shape, h, D, dt, K) as b:
        a.set_field(c_old)
        a.set_uptake(q)
        a.step("fp64", tol=1e-12, max_iter=200)
        c1 = a.field()
        with pytest.raises(BcgError):
            adj.adjoint(a, 2 * (c1 - target), "fp64", tol=1e-14, max_iter=1)
        assert a.field().tobytes() == c1.tobytes()
        a.step("fp64", tol=1e-12, max_iter=200)
        b.set_field(c1)
        b.set_uptake(q)
        b.step("fp64", tol=1e-12, max_iter=200)
        assert a.field().tobytes() == b.field().tobytes()


@pytest.mark.gpu
def test_mixed_adjoint_after_fp64_step_rebuilds_fp32_operator():
    """An FP64 step's set-up writes no FP32 copies of a0; a mixed-precision
    adjoint after it makes them and gives the gradients of a set-up that
    wrote them (set_state from the same c_old and c_new), bitwise."""
    from baccolonygrowth_amd import adjoint as adj

    shape = (9, 17, 17)
    h, D, dt, K = 1.0, 20.0, 10.0, 0.1
    c_old, q, target = _loss_parts(shape, seed=4)
    print(f"[gpu] mixed adjoint after an FP64 step {shape}", flush=True)
    with _solver(shape, h, D, dt, K) as a, _solver(shape, h, D, dt, K) as b:
        a.set_field(c_old)
        a.set_uptake(q)
        a.step("fp64", tol=1e-12, max_iter=200)
        c1 = a.field()
        g = 2 * (c1 - target)
        ra = adj.adjoint(a, g, "mixed", tol=1e-12)
        b.set_uptake(q)
        adj.set_state(b, c_old, c1)
        rb = adj.adjoint(b, g, "mixed", tol=1e-12)
    assert ra[0].tobytes() == rb[0].tobytes() and ra[1].tobytes() == rb[1].tobytes()
    assert ra[2] == rb[2] and ra[3] == rb[3]


@pytest.mark.gpu
def test_autograd_rates_D_K_vs_finite_differences():
    """Gradients of a radial-profile loss with respect to per-cell rates, D
    and K through torch.autograd, against central differences (h = 1e-6
    relative) of GPU FP64 solves."""
    import torch

    from baccolonygrowth_amd.adjoint import nutrient_step

    shape = (9, 33, 33)
    h, dt = 2.0, 10.0
    rng = np.random.default_rng(4)
    n = 3000
    cm = np.column_stack([rng.uniform(8, 56, n), rng.uniform(8, 56, n), rng.uniform(0.5, 8, n)])
    base = np.exp(rng.normal(0, 0.3, n)) * 2.0
    zz, yy, xx = np.meshgrid(np.arange(shape[0]) * h, np.arange(shape[1]) * h, np.arange(shape[2]) * h, indexing="ij")
    r = np.hypot(xx - 32, yy - 32)
    target = torch.from_numpy(20.0 * (1 - 0.8 * np.exp(-r / 150.0)))
    c0 = torch.full(shape, 20.0, dtype=torch.float64)
    with _solver(shape, h, 600.0, dt, 0.1) as s:
        from baccolonygrowth_amd.adjoint import set_coeffs

        def loss(rates, D, K):
            c = nutrient_step(rates, D, K, c0, s, cm, "fp64", 1e-12)
            return ((c - target) ** 2).sum()

        def loss_fd(rates_np, D, K):
            # FP64 solve run to the round-off floor (fixed V-cycle count)
            set_coeffs(s, D, K)
            s.deposit(cm, rates_np)
            s.set_field(c0.numpy())
            s.step("fp64", tol=1e-16, max_iter=40, strict=False)
            return float(((s.field() - target.numpy()) ** 2).sum())

        rates = torch.tensor(base, requires_grad=True)
        D = torch.tensor(600.0, dtype=torch.float64, requires_grad=True)
        K = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
        L0 = loss(rates, D, K)
        L0.backward()
        e = 1e-6
        for name, grad in (("D", D.grad.item()), ("K", K.grad.item())):
            v = {"D": 600.0, "K": 0.1}
            vp, vm = dict(v), dict(v)
            vp[name] *= 1 + e
            vm[name] *= 1 - e
            fd = (loss_fd(base, **vp) - loss_fd(base, **vm)) / (2 * e * v[name])
            print(f"    d/d{name}: adjoint {grad:.8e}  fd {fd:.8e}", flush=True)
            assert grad == pytest.approx(fd, rel=1e-4)
        for i in (0, 17, 1234):
            rp, rm = base.copy(), base.copy()
            rp[i] *= 1 + e
            rm[i] *= 1 - e
            fd = (loss_fd(rp, 600.0, 0.1) - loss_fd(rm, 600.0, 0.1)) / (2 * e * base[i])
            g = rates.grad[i].item()
            print(f"    d/drate[{i}]: adjoint {g:.8e}  fd {fd:.8e}", flush=True)
            assert g == pytest.approx(fd, rel=1e-3, abs=1e-6 * np.abs(rates.grad.numpy()).max())


@pytest.mark.gpu
def test_rollout_checkpointed_equals_direct_chain():
    """4 steps with checkpoints every 2 (recompute in backward) give the
    same gradients as the oracle's step-by-step adjoint chain."""
    import torch

    from baccolonygrowth_amd.adjoint import rollout

    shape = (5, 9, 9)
    h, D, dt, K = 1.0, 5.0, 10.0, 0.1
    rng = np.random.default_rng(8)
    n = 200
    cm = np.column_stack([rng.uniform(1, 7, n), rng.uniform(1, 7, n), rng.uniform(0.5, 3.5, n)])
    base = rng.random(n) * 0.5
    target = 20.0 * (1 - 0.2 * rng.random(shape))
    c0 = np.full(shape, 20.0)
    # oracle chain
    q, _ = rd3d.deposit_cic(cm, base, shape, h, (0, 0, 0))
    cs = [c0]
    for _ in range(4):
        cs.append(rd3d.step_direct(cs[-1], q, K, D, dt, h, 20.0))
    g = 2 * (cs[-1] - target)
    dq_tot = np.zeros(shape)
    dK_tot = dD_tot = 0.0
    for k in range(4, 0, -1):
        dq, dc0, dK, dD = rd3d.adjoint_direct(cs[k - 1], cs[k], q, K, D, dt, h, g)
        dq_tot += dq
        dK_tot += dK
        dD_tot += dD
        g = dc0
    with _solver(shape, h, D, dt, K) as s:
        from baccolonygrowth_amd.adjoint import gather

        rates = torch.tensor(base, requires_grad=True)
        Dt = torch.tensor(D, dtype=torch.float64, requires_grad=True)
        Kt = torch.tensor(K, dtype=torch.float64, requires_grad=True)
        c = rollout(rates, Dt, Kt, torch.from_numpy(c0), s, cm, steps=4, checkpoint_every=2)
        ((c - torch.from_numpy(target)) ** 2).sum().backward()
        drates_ref = gather(s, cm, dq_tot)
    assert np.abs(c.detach().numpy() - cs[-1]).max() <= 1e-9
    assert Kt.grad.item() == pytest.approx(dK_tot, rel=1e-7)
    assert Dt.grad.item() == pytest.approx(dD_tot, rel=1e-7)
    assert np.abs(rates.grad.numpy() - drates_ref).max() <= 1e-7 * np.abs(drates_ref).max()


@pytest.mark.gpu
def test_rollout_with_sources_per_step_equals_direct_chain():
    """4 steps whose sources change every step (cells that move), with
    checkpoints every 3 (a partial last window), give the gradients of the
    oracle's step-by-step adjoint chain: D, K and each step's rates."""
    import torch

    from baccolonygrowth_amd.adjoint import gather, rollout_seq

    shape = (5, 9, 9)
    h, D, dt, K = 1.0, 5.0, 10.0, 0.1
    rng = np.random.default_rng(9)
    n = 150
    cms = [np.column_stack([rng.uniform(1, 7, n), rng.uniform(1, 7, n), rng.uniform(0.5, 3.5, n)]) for _ in range(4)]
    bases = [rng.random(n) * 0.5 for _ in range(4)]
    target = 20.0 * (1 - 0.2 * rng.random(shape))
    c0 = np.full(shape, 20.0)
    qs = [rd3d.deposit_cic(cms[k], bases[k], shape, h, (0, 0, 0))[0] for k in range(4)]
    cs = [c0]
    for k in range(4):
        cs.append(rd3d.step_direct(cs[-1], qs[k], K, D, dt, h, 20.0))
    g = 2 * (cs[-1] - target)
    dqs = [None] * 4
    dK_tot = dD_tot = 0.0
    for k in range(3, -1, -1):
        dq, dc0, dK, dD = rd3d.adjoint_direct(cs[k], cs[k + 1], qs[k], K, D, dt, h, g)
        dqs[k] = dq
        dK_tot += dK
        dD_tot += dD
        g = dc0
    print("[gpu] rollout with per-step sources", flush=True)
    with _solver(shape, h, D, dt, K) as s:
        rates = [torch.tensor(b, requires_grad=True) for b in bases]
        Dt = torch.tensor(D, dtype=torch.float64, requires_grad=True)
        Kt = torch.tensor(K, dtype=torch.float64, requires_grad=True)
        c = rollout_seq(rates, Dt, Kt, torch.from_numpy(c0), s, cms, checkpoint_every=3)
        ((c - torch.from_numpy(target)) ** 2).sum().backward()
        refs = [gather(s, cms[k], dqs[k]) for k in range(4)]
    assert np.abs(c.detach().numpy() - cs[-1]).max() <= 1e-9
    assert Kt.grad.item() == pytest.approx(dK_tot, rel=1e-7)
    assert Dt.grad.item() == pytest.approx(dD_tot, rel=1e-7)
    for k in range(4):
        assert np.abs(rates[k].grad.numpy() - refs[k]).max() <= 1e-7 * np.abs(refs[k]).max()


@pytest.mark.gpu
def test_coupling_to_a_mechanics_context():
    """bcg_rd3d_deposit_cells equals bcg_rd3d_deposit of the same sources
    (centre (p + q) / 2, rate * Length), bitwise; bcg_rd3d_growth_to_cells
    gives mu c / (K + c) at every bin centre (trilinear in the field, clamped
    to the grid); coupled steps (mechanics, deposition, one FP64 V-cycle,
    growth rates) run on one device without host copies of the cells."""
    from baccolonygrowth_amd.colony import dome_colony, dome_params
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params
    from baccolonygrowth_amd.simulation import Colony

    print("[gpu] rd3d coupled to a 20000-cell colony", flush=True)
    cells = dome_colony(20000, radius=110.0, height=12.0, seed=5)
    p = dome_params(cells, bin_size=5.0, max_neighbours=96, extra_cells=20000)
    nz, ny, nx, h = 17, 65, 65, 2.0
    origin = (-64.0, -64.0, 0.0)
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=h, D=600.0, dt=10.0, K=0.1, c_res=20.0, origin=origin)
    rate, mu, K = 0.05, 1.2, 0.1
    with Colony(p) as col, Nutrients3D(prm) as s:
        col.LoadCells(cells)
        inside = s.deposit_cells(col, rate, count=True)
        q_dev = s.uptake()
        cm = (cells["p"] + cells["q"]) * 0.5
        assert s.deposit(cm, rate * cells["Length"]) == inside and 0 < inside < len(cells)
        assert q_dev.tobytes() == s.uptake().tobytes()

        f = 20.0 * np.random.default_rng(1).random((nz, ny, nx))
        s.set_field(f)
        s.growth_to_cells(col, mu, K)
        gr1 = col.fields()[3]
        b = np.arange(p.BoxX * p.BoxY * p.BoxZ)
        xb, yb, zb = b // (p.BoxY * p.BoxZ), (b // p.BoxZ) % p.BoxY, b % p.BoxZ
        pos = [(xb - p.BoxX // 2 + 0.5) * p.BoxLength, (yb - p.BoxY // 2) * p.BoxLength, (zb + 0.5) * p.BoxLength]
        idx, w = [], []
        for g, n in zip([(pos[a] - origin[a]) / h for a in range(3)], (nx, ny, nz)):
            g = np.clip(g, 0.0, n - 1)
            i0 = np.minimum(g.astype(int), n - 2)
            idx.append(i0)
            w.append(g - i0)
        (i, j, k), (wx, wy, wz) = idx, w
        c = sum(f[k + dk, j + dj, i + di] * (wx if di else 1 - wx) * (wy if dj else 1 - wy) * (wz if dk else 1 - wz)
                for di in (0, 1) for dj in (0, 1) for dk in (0, 1))
        c = np.maximum(c, 0.0)
        expect = mu * c / (K + c)
        assert np.abs(gr1 - expect).max() <= 1e-12 * expect.max()

        s.set_field(np.full((nz, ny, nx), 20.0))
        n0 = len(col.cells())
        for step in range(3):
            col.step(1.0 / 3600, step / 3600, update=True)
            s.deposit_cells(col, rate)
            s.vcycle("fp64")
            s.growth_to_cells(col, mu, K)
        field = s.field()
        assert len(col.cells()) >= n0 and np.isfinite(field).all() and field.min() < 20.0
        assert np.isfinite(col.fields()[3]).all() and col.fields()[3].min() < mu * 20.0 / (K + 20.0)


@pytest.mark.gpu
def test_coupled_update_graph_equals_separate_calls():
    """bcg_rd3d_update_cells (deposition, then set-up, one FP64 V-cycle and
    growth rates replayed from a recorded HIP graph) gives the field, the
    growth rates and the cells of the separate calls (deposit_cells, vcycle,
    growth_to_cells), bitwise, over three coupled steps with mechanics steps
    in between; mu changes on the third (the graph is recorded again)."""
    from baccolonygrowth_amd.colony import dome_colony, dome_params
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params
    from baccolonygrowth_amd.simulation import Colony

    print("[gpu] rd3d coupled update: graph replay vs separate calls", flush=True)
    cells = dome_colony(20000, radius=110.0, height=12.0, seed=6)
    p = dome_params(cells, bin_size=5.0, max_neighbours=96, extra_cells=20000)
    nz, ny, nx, h = 17, 65, 65, 2.0
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=h, D=600.0, dt=1.0, K=0.1, c_res=20.0, origin=(-64.0, -64.0, 0.0))
    rate, K = 0.5, 0.1
    with Colony(p) as ca, Colony(p) as cb, Nutrients3D(prm) as sa, Nutrients3D(prm) as sb:
        for col in (ca, cb):
            col.LoadCells(cells)
        for step in range(3):
            mu = 1.2 if step < 2 else 0.9
            for col in (ca, cb):
                col.step(1.0 / 3600, step / 3600, update=True)
            sa.update_cells(ca, rate, mu, K, 1)
            sb.deposit_cells(cb, rate)
            sb.vcycle("fp64")
            sb.growth_to_cells(cb, mu, K)
            fa, fb = sa.field(), sb.field()
            assert fa.tobytes() == fb.tobytes() and fa.min() < 20.0
            assert ca.fields()[3].tobytes() == cb.fields()[3].tobytes()
        assert ca.cells().tobytes() == cb.cells().tobytes()


@pytest.mark.gpu
def test_overlapped_coupled_steps_equal_serialised():
    """Coupled steps queued back to back (each nutrient update runs alongside
    the next mechanics step's bins and lists, and that step's move/grow waits
    for the growth rates) give the cells, the growth rates and the field of
    the same steps with the host waiting for every update, bitwise."""
    from baccolonygrowth_amd.colony import dome_colony, dome_params
    from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params
    from baccolonygrowth_amd.simulation import Colony

    print("[gpu] rd3d coupled steps: overlapped vs serialised", flush=True)
    cells = dome_colony(100000, radius=190.0, height=20.0, seed=7)
    p = dome_params(cells, bin_size=5.0, max_neighbours=96, extra_cells=100000)
    nz, ny, nx, h = 17, 129, 129, 4.0
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=h, D=600.0, dt=1.0, K=0.1, c_res=20.0,
                              origin=(-256.0, -256.0, 0.0))
    rate, mu, K = 0.5, 1.2, 0.1
    with Colony(p) as ca, Colony(p) as cb, Nutrients3D(prm) as sa, Nutrients3D(prm) as sb:
        for col in (ca, cb):
            col.LoadCells(cells)
        for step in range(6):
            ca.step(1.0 / 3600, step / 3600, update=True)
            sa.update_cells(ca, rate, mu, K, 2)
        for step in range(6):
            cb.step(1.0 / 3600, step / 3600, update=True)
            sb.update_cells(cb, rate, mu, K, 2)
            sb.field()  # the host waits for the update
            cb.fields()
        fa, fb = sa.field(), sb.field()
        assert fa.tobytes() == fb.tobytes() and fa.min() < 20.0
        assert ca.fields()[3].tobytes() == cb.fields()[3].tobytes()
        assert len(ca.cells()) >= len(cells) and ca.cells().tobytes() == cb.cells().tobytes()


@pytest.mark.gpu
def test_difference_form_central_differences():
    """loss_difference (each step's field difference solved directly,
    bcg_rd3d_solve) equals the difference of separately solved losses where
    that is accurate (h = 1e-2), and at h = 1e-6 its central differences
    match the adjoint gradients of a 3-step chain with per-step sources
    (D, K, a global rate factor, one cell's rate) to 1e-7."""
    import torch

    from baccolonygrowth_amd.adjoint import loss_difference, rollout_seq, set_coeffs

    shape = (9, 17, 17)
    h, D0, dt, K0 = 2.0, 600.0, 10.0, 0.1
    rng = np.random.default_rng(12)
    n = 400
    cms = [np.column_stack([rng.uniform(4, 28, n), rng.uniform(4, 28, n), rng.uniform(0.5, 12, n)]) for _ in range(3)]
    bases = [rng.lognormal(0.0, 0.3, n) * 3.0 for _ in range(3)]
    target = 20.0 * (1 - 0.3 * rng.random(shape))
    c0 = np.full(shape, 20.0)
    print("[gpu] difference-form central differences", flush=True)
    with _solver(shape, h, D0, dt, K0) as s:
        def plain_loss(D, K, a, rates=None):
            set_coeffs(s, D, K)
            c = c0
            for k in range(3):
                s.deposit(cms[k], a * (bases[k] if rates is None or k != 1 else rates))
                s.set_field(c)
                s.step("fp64", tol=1e-16, max_iter=30, strict=False)
                c = s.field()
            return float(((c - target) ** 2).sum())

        def diff(name, eps, i=7):
            Dm, Dp, Km, Kp = D0, D0, K0, K0
            rm, dr = list(bases), [None, None, None]
            if name == "D":
                Dm, Dp = D0 * (1 - eps), D0 * (1 + eps)
            elif name == "K":
                Km, Kp = K0 * (1 - eps), K0 * (1 + eps)
            elif name == "a":
                rm = [b * (1 - eps) for b in bases]
                dr = [b * (2 * eps) for b in bases]
            else:  # one cell's rate in the second step
                r = bases[1].copy()
                r[i] *= 1 - eps
                rm[1] = r
                dr[1] = np.zeros(n)
                dr[1][i] = 2 * eps * bases[1][i]
            return loss_difference(s, cms, rm, dr, c0, target, Dm, Dp, Km, Kp)

        # h = 1e-2: both forms are accurate; they agree
        e = 1e-2
        for name, args in (("D", ((D0 * (1 + e), K0, 1.0), (D0 * (1 - e), K0, 1.0))),
                           ("K", ((D0, K0 * (1 + e), 1.0), (D0, K0 * (1 - e), 1.0))),
                           ("a", ((D0, K0, 1 + e), (D0, K0, 1 - e)))):
            ref = plain_loss(*args[0]) - plain_loss(*args[1])
            got = diff(name, e)
            assert got == pytest.approx(ref, rel=1e-8), name
        # the adjoint gradients of the same chain
        rates = [torch.tensor(b, requires_grad=True) for b in bases]
        Dt = torch.tensor(D0, dtype=torch.float64, requires_grad=True)
        Kt = torch.tensor(K0, dtype=torch.float64, requires_grad=True)
        c = rollout_seq(rates, Dt, Kt, torch.from_numpy(c0), s, cms, checkpoint_every=2, mode="fp64", tol=1e-12)
        ((c - torch.from_numpy(target)) ** 2).sum().backward()
        g = {"D": Dt.grad.item() * D0, "K": Kt.grad.item() * K0,
             "a": sum(float((r.grad * torch.from_numpy(b)).sum()) for r, b in zip(rates, bases)),
             "r": rates[1].grad[7].item() * bases[1][7]}
        e = 1e-6
        for name in ("D", "K", "a", "r"):
            fd = diff(name, e) / (2 * e)
            print(f"    {name}: adjoint {g[name]:.12e}  central difference (h 1e-6) {fd:.12e}", flush=True)
            assert fd == pytest.approx(g[name], rel=1e-7), name
