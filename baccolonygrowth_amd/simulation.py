"""Python mirror of the reference's mechanics call surface, on the GPU.

`Colony` owns one device context (libbcg.so) and exposes the reference's
operations under their names:

    Colony.LoadCells(cells)               InputOutput.cpp:394 (minus the file read)
    Colony.getNeighbours(ids=None)        Neighbours.cpp:24-55
    Colony.sum_forces()                   Forces.cpp:332 (all cells, diagnostic)
    Colony.step(dt, t, update=...)        Simulation.cpp:330-413 (MoveCell,
                                          GrowCell, DivideCell, list patch, swap)
    Colony.srand(seed) / rand()           the rand() stream the reference draws from

Errors come back as BcgError carrying the C-ABI code (the reference instead
exits through MyAssert, tools.h:18-27).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import BcgError, _dp, _ip, _vp, lib
from .cells import CELL_DTYPE, Params

STEP_REBUILD = 1
STEP_NO_PATCH = 2


def default_params() -> Params:
    p = Params()
    _check(None, lib().bcg_default_params(ctypes.byref(p)))
    return p


def _check(ctx, rc):
    if rc != 0:
        msg = ""
        if ctx:
            buf = ctypes.create_string_buffer(512)
            lib().bcg_last_error(ctx, buf, 512)
            msg = buf.value.decode(errors="replace")
        raise BcgError(rc, msg)


def _d(a):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(_dp), a


class Colony:
    def __init__(self, params: Params):
        self.params = params.copy()
        h = _vp()
        _check(None, lib().bcg_create(ctypes.byref(self.params), ctypes.byref(h)))
        self._h = h

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            lib().bcg_destroy(self._h)
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

    def _c(self, rc):
        _check(self._h, rc)

    # -- state
    def LoadCells(self, cells: np.ndarray):
        cells = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        self._c(lib().bcg_set_cells(self._h, cells.ctypes.data_as(_vp), len(cells)))

    set_cells = LoadCells

    @property
    def n(self) -> int:
        n = ctypes.c_int32()
        self._c(lib().bcg_num_cells(self._h, ctypes.byref(n)))
        return n.value

    def cells(self) -> np.ndarray:
        n = self.n
        out = np.zeros(n, CELL_DTYPE)
        got = ctypes.c_int32()
        self._c(lib().bcg_get_cells(self._h, out.ctypes.data_as(_vp), n, ctypes.byref(got)))
        return out

    def set_fields(self, height=None, normal=None, wall=None, gr1=None, gr2=None):
        ptrs = [_d(a) for a in (height, normal, wall, gr1, gr2)]
        self._c(lib().bcg_set_fields(self._h, *[p for p, _ in ptrs]))

    def fields(self):
        """(height, normal, wall, gr1, gr2) as the mechanics sees them."""
        p = self.params
        nh = p.BoxX * p.refinementGridHeight * p.BoxY * p.refinementGridHeight
        nb = p.BoxX * p.BoxY * p.BoxZ
        out = [np.zeros(nh), np.zeros(3 * nh), np.zeros(p.BoxX * p.BoxY), np.zeros(nb), np.zeros(nb)]
        self._c(lib().bcg_get_fields(self._h, *[a.ctypes.data_as(_dp) for a in out]))
        return tuple(out)

    # -- RNG
    def srand(self, seed: int):
        self._c(lib().bcg_srand(self._h, seed))

    def rand(self) -> int:
        v = ctypes.c_int32()
        self._c(lib().bcg_rand(self._h, ctypes.byref(v)))
        return v.value

    # -- neighbours / forces / step
    def getNeighbours(self, ids=None):
        if ids is None:
            self._c(lib().bcg_rebuild_neighbours(self._h, None, -1))
        else:
            ids = np.ascontiguousarray(ids, np.int32)
            self._c(lib().bcg_rebuild_neighbours(self._h, ids.ctypes.data_as(_ip), len(ids)))

    def neighbours(self, stride: int | None = None):
        """(counts[N], lists[N, stride]) with each list sorted ascending."""
        stride = stride or self.params.maxNeighbours
        n = self.n
        counts = np.zeros(n, np.int32)
        lists = np.full((n, stride), -1, np.int32)
        self._c(lib().bcg_get_neighbours(self._h, counts.ctypes.data_as(_ip), lists.ctypes.data_as(_ip), stride))
        return counts, lists

    def sum_forces(self):
        n = self.n
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        self._c(lib().bcg_sum_forces(self._h, F.ctypes.data_as(_dp), T.ctypes.data_as(_dp)))
        return F, T

    def mean_stress(self):
        """mean_stress of every cell as the reference's output step computes
        it: (stress[N, 5] = xx, yy, zz, rr, tt; Fnet[N, 3]). Takes the rand()
        draws of its F_cw calls."""
        n = self.n
        S = np.zeros((n, 5))
        F = np.zeros((n, 3))
        self._c(lib().bcg_mean_stress(self._h, S.ctypes.data_as(_dp), F.ctypes.data_as(_dp)))
        return S, F

    def step(self, dt: float, t: float, update: bool = False, no_patch: bool = False, lineage_cap: int = 1 << 16):
        """One mechanics step; returns the (parent, daughter) pairs."""
        flags = (STEP_REBUILD if update else 0) | (STEP_NO_PATCH if no_patch else 0)
        nd = ctypes.c_int32()
        lin = np.zeros((lineage_cap, 2), np.int32)
        self._c(lib().bcg_step(self._h, dt, t, flags, ctypes.byref(nd), lin.ctypes.data_as(_ip), lineage_cap))
        return lin[: min(nd.value, lineage_cap)].copy()

    def last_step_stats(self):
        w = ctypes.c_int64()
        f = ctypes.c_int32()
        m = ctypes.c_int32()
        self._c(lib().bcg_last_step_stats(self._h, ctypes.byref(w), ctypes.byref(f), ctypes.byref(m)))
        return {"wall_draws": w.value, "flagged": f.value, "max_list": m.value}

    def last_step_times(self):
        ms = (ctypes.c_float * 5)()
        self._c(lib().bcg_last_step_times(self._h, ms))
        return {"bins": ms[0], "lists": ms[1], "move_grow": ms[2], "compact": ms[3], "divide": ms[4]}


__all__ = ["Colony", "BcgError", "default_params", "STEP_REBUILD", "STEP_NO_PATCH"]
