"""Differentiable nutrient step: torch.autograd.Function over the C ABI.

`nutrient_step(rates, D, K, c_old, solver, cm)` runs one implicit 3-D step
(bcg_rd3d_step) with the uptake density deposited from per-cell rates at
centres cm, and returns c_new. Its backward solves the adjoint system with
the same multigrid (bcg_rd3d_adjoint) and returns gradients with respect to
the per-cell rates (transpose of the deposition, bcg_rd3d_gather), D, K and
c_old, so steps chain. `rollout(...)` applies several steps with the field
checkpointed every k steps; the steps in between are recomputed in the
backward pass (torch.utils.checkpoint).

Tensors are float64 on the host; the solver keeps its state in HBM.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import _dp, lib
from .nutrients import MODES, Nutrients3D


def _np(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().cpu().numpy(), dtype=np.float64)


def adjoint(solver: Nutrients3D, g: np.ndarray, mode: str = "fp64", tol: float = 1e-12, max_iter: int = 200):
    """(dL/dq, dL/dc_old, dL/dK, dL/dD) of the solver's last step."""
    g = solver._arr(g)
    dq = np.empty(solver.shape)
    dc = np.empty(solver.shape)
    dK = ctypes.c_double()
    dD = ctypes.c_double()
    solver._check(
        lib().bcg_rd3d_adjoint(
            solver._h,
            g.ctypes.data_as(_dp),
            MODES[mode],
            tol,
            max_iter,
            dq.ctypes.data_as(_dp),
            dc.ctypes.data_as(_dp),
            ctypes.byref(dK),
            ctypes.byref(dD),
        )
    )
    return dq, dc, dK.value, dD.value


def gather(solver: Nutrients3D, cm: np.ndarray, field: np.ndarray) -> np.ndarray:
    cm = np.ascontiguousarray(cm, dtype=np.float64).reshape(-1, 3)
    field = solver._arr(field)
    out = np.empty(len(cm))
    solver._check(
        lib().bcg_rd3d_gather(solver._h, cm.ctypes.data_as(_dp), len(cm), field.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    )
    return out


def set_state(solver: Nutrients3D, c_old: np.ndarray, c_new: np.ndarray) -> None:
    a, b = solver._arr(c_old), solver._arr(c_new)
    solver._check(lib().bcg_rd3d_set_state(solver._h, a.ctypes.data_as(_dp), b.ctypes.data_as(_dp)))


def set_coeffs(solver: Nutrients3D, D: float, K: float) -> None:
    solver._check(lib().bcg_rd3d_set_coeffs(solver._h, float(D), float(K)))


class NutrientStepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rates, D, K, c_old, solver, cm, mode, tol, adj_tol):
        set_coeffs(solver, D.item(), K.item())
        solver.deposit(cm, _np(rates))
        solver.set_field(_np(c_old))
        solver.step(mode, tol=tol, max_iter=200)
        c_new = torch.from_numpy(solver.field())
        ctx.save_for_backward(rates, D, K, c_old, c_new)
        ctx.solver, ctx.cm, ctx.mode = solver, cm, mode
        ctx.adj_tol = adj_tol if adj_tol is not None else min(tol, 1e-12)
        return c_new

    @staticmethod
    def backward(ctx, g):
        rates, D, K, c_old, c_new = ctx.saved_tensors
        s = ctx.solver
        # restore this step's state (another step may have run since)
        set_coeffs(s, D.item(), K.item())
        s.deposit(ctx.cm, _np(rates))
        set_state(s, _np(c_old), _np(c_new))
        dq, dc, dK, dD = adjoint(s, _np(g), ctx.mode, ctx.adj_tol)
        drates = torch.from_numpy(gather(s, ctx.cm, dq))
        return (
            drates,
            torch.tensor(dD, dtype=torch.float64),
            torch.tensor(dK, dtype=torch.float64),
            torch.from_numpy(dc),
            None,
            None,
            None,
            None,
            None,
        )


def nutrient_step(rates, D, K, c_old, solver: Nutrients3D, cm, mode: str = "fp64", tol: float = 1e-12,
                  adj_tol: float | None = None):
    """One implicit step as a torch.autograd.Function; the adjoint solve of
    its backward pass goes to `adj_tol` (default min(tol, 1e-12))."""
    return NutrientStepFn.apply(rates, D, K, c_old, solver, cm, mode, tol, adj_tol)


def rollout(rates, D, K, c0, solver: Nutrients3D, cm, steps: int, checkpoint_every: int = 5, mode="fp64", tol=1e-12):
    """`steps` nutrient steps; only every `checkpoint_every`-th field is kept
    for the backward pass, the others are recomputed from it."""
    from torch.utils.checkpoint import checkpoint

    def window(c, n):
        for _ in range(n):
            c = nutrient_step(rates, D, K, c, solver, cm, mode, tol)
        return c

    c = c0
    done = 0
    while done < steps:
        n = min(checkpoint_every, steps - done)
        c = checkpoint(window, c, n, use_reentrant=False)
        done += n
    return c


def rollout_seq(rates_seq, D, K, c0, solver: Nutrients3D, cm_seq, checkpoint_every: int = 5, mode="fp64", tol=1e-12,
                adj_tol=None):
    """One nutrient step per (rates, cm) pair of the sequences: the sources of
    each update (the cells move and divide between updates); only every
    `checkpoint_every`-th field is kept for the backward pass, the others are
    recomputed from it."""
    from torch.utils.checkpoint import checkpoint

    def window(c, lo, hi):
        for k in range(lo, hi):
            c = nutrient_step(rates_seq[k], D, K, c, solver, cm_seq[k], mode, tol, adj_tol)
        return c

    c = c0
    n = len(rates_seq)
    for lo in range(0, n, checkpoint_every):
        c = checkpoint(window, c, lo, min(lo + checkpoint_every, n), use_reentrant=False)
    return c


def laplacian_h2(u: np.ndarray) -> np.ndarray:
    """h^2 times the 7-point Laplacian of the step's operator on [z][y][x]
    arrays: mirror ghosts on the x and y faces and on the top face (rows of
    the bottom face are Dirichlet; callers set them themselves)."""
    p = np.pad(u, 1, mode="reflect")
    out = p[1:-1, 1:-1, 2:] + p[1:-1, 1:-1, :-2]
    out += p[1:-1, 2:, 1:-1]
    out += p[1:-1, :-2, 1:-1]
    out += p[2:, 1:-1, 1:-1]
    out += p[:-2, 1:-1, 1:-1]
    out -= 6.0 * u
    return out


def loss_difference(solver: Nutrients3D, cm_seq, rates_minus, drates, c0, target, D_minus: float, D_plus: float,
                    K_minus: float, K_plus: float, max_iter: int = 30) -> float:
    """L(+) - L(-) for L = sum (c_n - target)^2, c_n the field after one
    implicit FP64 step per source set (cm_seq[k], rates), from c0; (-) has
    diffusivity D_minus, Monod constant K_minus and rates rates_minus[k], (+)
    has D_plus, K_plus and rates_minus[k] + drates[k] (drates[k] may be None).

    Evaluated in difference form: with A-/A+ the two steps' operators, the
    difference d = c+ - c- of every step solves exactly
        A+_k d_k = d_{k-1} + (A-_k - A+_k) c-_k,
        A-_k - A+_k = diag(a0-_k - a0+_k) + (lam+ - lam-) Lap_h,
    a0 = 1 + dt q / (K + c_prev), lam = dt D / h^2, with a0- - a0+ formed from
    the differences (dq, dK, d_{k-1}) directly. Subtracting two separately
    solved fields instead loses the difference below the solves' round-off
    floor (the operator's terms are ~lam |c|, lam ~ 1e3); here each d_k is
    solved to round-off relative to itself, so a central difference at
    h = 1e-6 is as accurate as at h = 1e-2, less the h^2 truncation."""
    prm = solver.params
    dt, h = prm.dt, prm.h
    dlam = dt * (D_plus - D_minus) / (h * h)
    dK = K_plus - K_minus
    dn = np.ascontiguousarray(c0, dtype=np.float64)
    d = np.zeros_like(dn)
    for k in range(len(cm_seq)):
        cm = cm_seq[k]
        # (-) step from dn
        set_coeffs(solver, D_minus, K_minus)
        solver.deposit(cm, rates_minus[k])
        q_minus = solver.uptake()
        solver.set_field(dn)
        solver.step("fp64", tol=1e-16, max_iter=max_iter, strict=False)
        dn_new = solver.field()
        if drates is not None and drates[k] is not None:
            solver.deposit(cm, drates[k])
            dq = solver.uptake()
            rates_plus = rates_minus[k] + drates[k]
        else:
            dq = 0.0
            rates_plus = rates_minus[k]
        up_prev = dn + d
        km = K_minus + dn
        da0 = dt * (q_minus * (dK + d) - dq * km) / (km * (K_plus + up_prev))
        rhs = d + da0 * dn_new + dlam * laplacian_h2(dn_new)
        rhs[0] = 0.0
        # (+) operator: its coefficients, sources and c_prev = up_prev
        set_coeffs(solver, D_plus, K_plus)
        solver.deposit(cm, rates_plus)
        set_state(solver, up_prev, dn_new)
        d, _, _ = solver.solve(rhs, "fp64", tol=1e-16, max_iter=max_iter, strict=False)
        dn = dn_new
    t = np.asarray(target, dtype=np.float64)
    return float((d * (2.0 * dn + d - 2.0 * t)).sum())
