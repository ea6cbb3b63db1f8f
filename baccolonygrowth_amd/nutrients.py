"""3-D nutrient step on the GPU (bcg_rd3d_* of include/bcg.h).

`Nutrients3D` holds one nutrient grid and solves the implicit step
    (1 + dt q/(K + c_old)) c_new - dt D Lap_h c_new = c_old
(Neumann x/y/top faces, Dirichlet c_res at z = 0) by FP64 multigrid
(mode "fp64") or by FP32 V-cycles inside FP64 iterative refinement
(mode "mixed"). It is the 3-D counterpart of the reference's planar
nutrient update (Multigrid.cpp:984-1125); q comes from the cells by
cloud-in-cell deposition (the role of GetDensity, Compute.cpp:120-253).
Arrays are [nz][ny][nx] float64 on the host.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import RD3DParams, BcgError, _dp, _fp, _ip, _vp, lib

MODES = {"fp64": 0, "mixed": 1}


# bcg_rd3d_params.smoother (include/bcg.h)
SMOOTH_FUSED = 0  # both colours of a red-black sweep in one z-marching launch
SMOOTH_SPLIT = 1  # one launch per colour
SMOOTH_FUSED_WIDE = 2  # FUSED, with the FP64 sweeps of levels >= 2^24 points on the wide tile too


def default_rd3d_params(**kw) -> RD3DParams:
    p = RD3DParams()
    rc = lib().bcg_rd3d_default_params(ctypes.byref(p))
    if rc:
        raise BcgError(rc)
    for k, v in kw.items():
        if k == "origin":
            for i in range(3):
                p.origin[i] = v[i]
        else:
            setattr(p, k, v)
    return p


class Nutrients3D:
    def __init__(self, params: RD3DParams):
        self.params = params
        h = _vp()
        self._check(lib().bcg_rd3d_create(ctypes.byref(params), ctypes.byref(h)), None)
        self._h = h
        self.shape = (params.nz, params.ny, params.nx)

    def _check(self, rc, h="self"):
        if rc:
            msg = ""
            hh = self._h if h == "self" else h
            if hh:
                buf = ctypes.create_string_buffer(512)
                lib().bcg_rd3d_last_error(hh, buf, 512)
                msg = buf.value.decode(errors="replace")
            raise BcgError(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            lib().bcg_rd3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def levels(self):
        n = ctypes.c_int32()
        dims = np.zeros(3 * 32, np.int32)
        self._check(lib().bcg_rd3d_levels(self._h, ctypes.byref(n), dims.ctypes.data_as(_ip), 32))
        return [tuple(dims[3 * i : 3 * i + 3]) for i in range(n.value)]

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"expected {self.shape}, got {a.shape}")
        return a

    def set_field(self, c):
        c = self._arr(c)
        self._check(lib().bcg_rd3d_set_field(self._h, c.ctypes.data_as(_dp)))

    def field(self):
        out = np.empty(self.shape)
        self._check(lib().bcg_rd3d_get_field(self._h, out.ctypes.data_as(_dp)))
        return out

    def set_uptake(self, q):
        q = self._arr(q)
        self._check(lib().bcg_rd3d_set_uptake(self._h, q.ctypes.data_as(_dp)))

    def uptake(self):
        out = np.empty(self.shape)
        self._check(lib().bcg_rd3d_get_uptake(self._h, out.ctypes.data_as(_dp)))
        return out

    def deposit(self, cm, rate) -> int:
        cm = np.ascontiguousarray(cm, dtype=np.float64).reshape(-1, 3)
        rate = np.ascontiguousarray(rate, dtype=np.float64).reshape(-1)
        inside = ctypes.c_int32()
        self._check(
            lib().bcg_rd3d_deposit(self._h, cm.ctypes.data_as(_dp), rate.ctypes.data_as(_dp), len(rate), ctypes.byref(inside))
        )
        return inside.value

    def deposit_cells(self, colony, rate: float, count: bool = False):
        """Uptake density from the cells of `colony` (a simulation.Colony on the
        same device), each depositing rate * Length at its centre; no host
        copies. Returns the number of cells inside the grid when `count`
        (which synchronises), else None."""
        inside = ctypes.c_int32()
        self._check(lib().bcg_rd3d_deposit_cells(self._h, colony._h, float(rate), ctypes.byref(inside) if count else None))
        return inside.value if count else None

    def growth_to_cells(self, colony, mu: float, K: float) -> None:
        """Growth rate of every bin of `colony` = mu c / (K + c) at the bin
        centre, from the current field (the role of the reference's growth-rate
        lookup); ordered before the colony's next step."""
        self._check(lib().bcg_rd3d_growth_to_cells(self._h, colony._h, float(mu), float(K)))

    def update_cells(self, colony, rate: float, mu: float, K: float, vcycles: int = 1) -> None:
        """One coupled update without host synchronisation: deposition from the
        cells of `colony`, the implicit step by `vcycles` FP64 V-cycles, the
        growth rates of its bins (a replayed HIP graph for the fixed part)."""
        self._check(lib().bcg_rd3d_update_cells(self._h, colony._h, float(rate), float(mu), float(K), int(vcycles)))

    def step(self, mode: str = "mixed", tol: float = 1e-10, max_iter: int = 100, strict: bool = True):
        """One implicit step; returns (iterations, ||b - A c|| / ||b||).
        strict=False keeps the last iterate when tol is not reached (for
        solves run to the round-off floor)."""
        it = ctypes.c_int32()
        rr = ctypes.c_double()
        rc = lib().bcg_rd3d_step(self._h, MODES[mode], tol, max_iter, ctypes.byref(it), ctypes.byref(rr))
        if rc == -7 and not strict:
            return it.value, rr.value
        self._check(rc)
        return it.value, rr.value

    def solve(self, rhs, mode: str = "fp64", tol: float = 1e-12, max_iter: int = 100, strict: bool = True):
        """x with A x = rhs for the operator of the last step (or set_state),
        from x = 0; returns (x, iterations, ||rhs - A x|| / ||rhs||). The
        field and the step's right-hand side are unchanged. strict=False
        keeps the last iterate when tol is not reached."""
        rhs = self._arr(rhs)
        x = np.empty(self.shape)
        it = ctypes.c_int32()
        rr = ctypes.c_double()
        rc = lib().bcg_rd3d_solve(self._h, MODES[mode], rhs.ctypes.data_as(_dp), tol, max_iter, x.ctypes.data_as(_dp),
                                  ctypes.byref(it), ctypes.byref(rr))
        if not (rc == -7 and not strict):
            self._check(rc)
        return x, it.value, rr.value

    def vcycle(self, mode: str = "fp64") -> float:
        ms = ctypes.c_float()
        self._check(lib().bcg_rd3d_vcycle(self._h, MODES[mode], ctypes.byref(ms)))
        return ms.value

    def last_times(self):
        ms = (ctypes.c_float * 2)()
        self._check(lib().bcg_rd3d_last_times(self._h, ms))
        return {"setup": ms[0], "solve": ms[1]}
