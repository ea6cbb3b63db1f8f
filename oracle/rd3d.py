"""FP64 oracle of the 3-D implicit reaction-diffusion (nutrient) step.

TEST INFRASTRUCTURE ONLY: the checker for the HIP multigrid solver and its
adjoint. The reference solves its nutrient field only in a planar x-z model
(Multigrid.cpp / Nutrients.cpp); the 3-D step below is the generalisation
the 3-D benchmark configurations need, so this file DEFINES the discrete
problem and is pinned by mathematics, not by reference vectors:
  * a direct sparse solve (scipy.sparse.linalg.spsolve) on small grids;
  * the operator's invariants: constants are in the kernel of the
    diffusion part; the symmetrised operator W A is symmetric;
  * manufactured solutions (b := A u* for a chosen u*);
  * the conjugate-gradient form of the step and of its adjoint (for grids
    where the factorisation is slow) against the direct solves;
  * the cloud-in-cell deposition by conservation of the total rate and of
    its first moment (exact for linear weights).

The step, on a vertex grid of n = (nx, ny, nz) points with spacing h:
    (1 + dt*sigma) c_new - lam * Lap_h c_new = c_old,   lam = dt*D/h^2
    sigma = q / (K + c_old)          Monod uptake linearised about c_old
with Lap_h the 7-point Laplacian, Neumann (mirror) on the x, y faces and on
the top face z = nz-1, and Dirichlet c = c_res on the bottom face z = 0 (the
agar reservoir). q is the uptake-rate density deposited from the cells.
Arrays are indexed [z][y][x] (x fastest).
"""
from __future__ import annotations

import numpy as np


def laplacian(u: np.ndarray) -> np.ndarray:
    """7-point Laplacian (times h^2) with mirror ghosts on every face."""
    p = np.pad(u, 1, mode="reflect")
    return (
        p[1:-1, 1:-1, 2:]
        + p[1:-1, 1:-1, :-2]
        + p[1:-1, 2:, 1:-1]
        + p[1:-1, :-2, 1:-1]
        + p[2:, 1:-1, 1:-1]
        + p[:-2, 1:-1, 1:-1]
        - 6.0 * u
    )


def apply_A(u: np.ndarray, sigma: np.ndarray, lam: float, dt: float) -> np.ndarray:
    """A u with the Dirichlet rows (z = 0) as identity."""
    out = (1.0 + dt * sigma) * u - lam * laplacian(u)
    out[0] = u[0]
    return out


def residual(b, u, sigma, lam, dt):
    return b - apply_A(u, sigma, lam, dt)


def rhs(c_old: np.ndarray, c_res: float) -> np.ndarray:
    b = c_old.copy()
    b[0] = c_res
    return b


def sparse_A(shape, sigma, lam, dt):
    """The same operator as a scipy CSR matrix (rows in [z][y][x] order)."""
    import scipy.sparse as sp

    nz, ny, nx = shape
    N = nx * ny * nz
    idx = np.arange(N).reshape(shape)
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r.ravel())
        cols.append(c.ravel())
        vals.append(np.broadcast_to(v, r.shape).ravel())

    add(idx, idx, 1.0 + dt * sigma + 6.0 * lam)
    for axis, n in ((2, nx), (1, ny), (0, nz)):
        for s in (1, -1):
            nb = np.roll(np.arange(n), -s)
            if s == 1:
                nb[-1] = n - 2  # mirror at the high face
            else:
                nb[0] = 1  # mirror at the low face
            take = [slice(None)] * 3
            src = idx
            dst = np.take(idx, nb, axis=axis)
            add(src, dst, -lam)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    # Dirichlet rows
    A = A.tolil()
    for r in idx[0].ravel():
        A.rows[r] = [r]
        A.data[r] = [1.0]
    return A.tocsr()


def solve_direct(b, sigma, lam, dt):
    from scipy.sparse.linalg import spsolve

    A = sparse_A(b.shape, sigma, lam, dt)
    return spsolve(A.tocsc(), b.ravel()).reshape(b.shape)


def control_volume_weights(shape) -> np.ndarray:
    """W: 1/2 per Neumann face a vertex lies on (the symmetriser of A)."""
    nz, ny, nx = shape
    w = np.ones(shape)
    w[:, :, 0] *= 0.5
    w[:, :, -1] *= 0.5
    w[:, 0, :] *= 0.5
    w[:, -1, :] *= 0.5
    w[-1, :, :] *= 0.5
    return w


def deposit_cic(cm: np.ndarray, q: np.ndarray, shape, h: float, origin) -> np.ndarray:
    """Cloud-in-cell deposition of per-cell rates q at centres cm onto the
    vertex grid, divided by the vertex volume h^3 (rate density). Cells
    outside the grid are dropped. Returns [z][y][x]."""
    nz, ny, nx = shape
    g = (cm - np.asarray(origin)[None, :]) / h
    i0 = np.floor(g).astype(np.int64)
    f = g - i0
    out = np.zeros(shape)
    inside = (i0[:, 0] >= 0) & (i0[:, 0] < nx - 1) & (i0[:, 1] >= 0) & (i0[:, 1] < ny - 1) & (i0[:, 2] >= 0) & (i0[:, 2] < nz - 1)
    i0, f, qq = i0[inside], f[inside], q[inside]
    for dz in (0, 1):
        wz = f[:, 2] if dz else 1 - f[:, 2]
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1 - f[:, 1]
            for dx in (0, 1):
                wx = f[:, 0] if dx else 1 - f[:, 0]
                np.add.at(out, (i0[:, 2] + dz, i0[:, 1] + dy, i0[:, 0] + dx), qq * wx * wy * wz)
    return out / h**3, inside


def _free_system(shape, sigma, lam, dt):
    """(S, w, free, A): the rows of A that are not Dirichlet, restricted to
    the free unknowns and scaled by the control-volume weights w, S = W_f
    A_ff, which is symmetric positive definite (symmetric by construction of
    W; positive diagonal dominance from 1 + dt sigma > 0)."""
    A = sparse_A(shape, sigma, lam, dt)
    free = np.ones(shape, bool)
    free[0] = False
    f = free.ravel()
    w = control_volume_weights(shape).ravel()[f]
    S = (A[f][:, f]).multiply(w[:, None]).tocsr()
    return S, w, f, A


def solve_cg(b, sigma, lam, dt, rtol=1e-14):
    """The same system as solve_direct by conjugate gradients on the
    symmetrised free block (for grids where the sparse factorisation is too
    slow); Dirichlet values are b there."""
    from scipy.sparse.linalg import cg

    S, w, f, A = _free_system(b.shape, sigma, lam, dt)
    x = b.ravel().copy()
    fixed = x.copy()
    fixed[f] = 0.0
    rhs_f = w * (b.ravel()[f] - (A @ fixed)[f])
    xf, info = cg(S, rhs_f, x0=b.ravel()[f], rtol=rtol, maxiter=20000)
    if info != 0:
        raise RuntimeError(f"cg did not converge (info={info})")
    x[f] = xf
    return x.reshape(b.shape)


def step_direct(c_old, q, K, D, dt, h, c_res, solver="direct"):
    """One implicit step by direct solve (or solver="cg": conjugate
    gradients to 1e-14): returns c_new."""
    sigma = q / (K + c_old)
    if solver == "cg":
        return solve_cg(rhs(c_old, c_res), sigma, dt * D / h**2, dt)
    return solve_direct(rhs(c_old, c_res), sigma, dt * D / h**2, dt)


def adjoint_direct(c_old, c_new, q, K, D, dt, h, g, solver="direct"):
    """Gradients of L with dL/dc_new = g through one step, from the sparse
    transpose solve A^T lam = g (no symmetrisation used; solver="cg": the
    free block of the transpose through the symmetrised system,
    A_ff^T = S W_f^-1, by conjugate gradients to 1e-14):
    returns dL/dq, dL/dc_old, dL/dK, dL/dD."""
    from scipy.sparse.linalg import cg, spsolve

    shape = c_old.shape
    sigma = q / (K + c_old)
    lam_c = dt * D / h**2
    if solver == "cg":
        # only the free rows of lam enter the gradients below
        S, w, f, _ = _free_system(shape, sigma, lam_c, dt)
        mu, info = cg(S, g.ravel()[f], rtol=1e-14, maxiter=20000)
        if info != 0:
            raise RuntimeError(f"cg did not converge (info={info})")
        lam = np.zeros(g.size)
        lam[f] = w * mu
        lam = lam.reshape(shape)
    else:
        A = sparse_A(shape, sigma, lam_c, dt)
        lam = spsolve(A.T.tocsc(), g.ravel()).reshape(shape)
    free = np.ones(shape, bool)
    free[0] = False
    den = K + c_old
    ga0 = np.where(free, -lam * c_new, 0.0)
    da0_dc = -dt * q / den**2
    dq = ga0 * dt / den
    dc0 = np.where(free, lam + ga0 * da0_dc, 0.0)
    dK = float((ga0 * da0_dc).sum())
    dD = float((np.where(free, lam, 0.0) * laplacian(c_new)).sum() * dt / h**2)
    return dq, dc0, dK, dD
