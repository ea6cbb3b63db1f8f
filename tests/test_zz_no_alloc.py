"""bcg_rd3d_deposit_noalloc (no allocation, static device scratch) against
bcg_rd3d_deposit through the C ABI: the uptake density must be bit-identical,
on two calls in a row, the second with fewer sources (the scratch keeps the
first call's contents, so whatever must start at zero has to be cleared by
the call itself). The test's device buffers come from the HIP runtime
(hipMalloc through ctypes): the test allocates, the entry point does not."""
import ctypes
import os

import numpy as np
import pytest

from baccolonygrowth_amd._lib import BcgError, lib
from baccolonygrowth_amd.nutrients import Nutrients3D, default_rd3d_params

pytestmark = pytest.mark.gpu

SCRATCH = 16 << 20


def _sources(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    cm = rng.uniform(lo - 3.0, hi + 3.0, size=(n, 3))  # some outside the grid: dropped by both paths
    cm[: n // 7] = np.round(cm[: n // 7])  # some exactly on vertices and faces
    rate = rng.lognormal(0.0, 0.5, n)
    return np.ascontiguousarray(cm), np.ascontiguousarray(rate)


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime, for the test's own
    device buffers."""

    H2D, D2H = 1, 2

    def __init__(self):
        # the runtime libbcg.so is bound to (by the soname it names, after
        # loading it): a bare "libamdhip64.so" can resolve to another copy
        # that a library loaded earlier in the process brought in (torch's)
        lib()
        self.L = ctypes.CDLL("libamdhip64.so.7", mode=getattr(os, "RTLD_NOLOAD", 4) | ctypes.RTLD_GLOBAL)
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.live = []

    def put(self, a: np.ndarray) -> ctypes.c_void_p:
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p

    def get(self, p: ctypes.c_void_p, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def _noalloc(s, cm, rate, q_init=np.nan):
    hip = _Hip()
    try:
        q0 = np.full(s.shape, q_init)
        d_cm, d_rate, d_q = hip.put(cm), hip.put(rate), hip.put(q0)
        rc = lib().bcg_rd3d_deposit_noalloc(s._h, d_cm if len(rate) else None, d_rate if len(rate) else None,
                                            len(rate), d_q, None)
        s._check(rc)
        return hip.get(d_q, q0)
    finally:
        hip.free()


@pytest.mark.parametrize("shape,h", [((17, 65, 65), 2.0), ((33, 129, 129), 1.0)])
def test_noalloc_deposit_equals_deposit_twice(shape, h):
    nz, ny, nx = shape
    half = (nx - 1) * h / 2
    prm = default_rd3d_params(nx=nx, ny=ny, nz=nz, h=h, origin=(-half, -half, 0.0), max_levels=4)
    with Nutrients3D(prm) as s:
        hi = np.array([half, half, (nz - 1) * h])
        lo = np.array([-half, -half, 0.0])
        for n, seed in ((60000, 1), (7001, 2)):  # second call smaller than the first
            cm, rate = _sources(n, lo, hi, seed)
            inside = s.deposit(cm, rate)
            assert 0 < inside < n
            ref = s.uptake()
            got = _noalloc(s, cm, rate)
            assert np.array_equal(got, ref), f"n={n}: max |diff| {np.abs(got - ref).max():.3e}"
            # the context's own uptake field is left as the allocating call set it
            assert np.array_equal(s.uptake(), ref)


def test_noalloc_deposit_of_no_sources_clears():
    prm = default_rd3d_params(nx=33, ny=33, nz=9, h=2.0, origin=(-32.0, -32.0, 0.0), max_levels=3)
    with Nutrients3D(prm) as s:
        cm, rate = _sources(5000, np.array([-32.0, -32.0, 0.0]), np.array([32.0, 32.0, 16.0]), 3)
        assert _noalloc(s, cm, rate).sum() > 0
        # no sources after a call that left counts and keys in the scratch
        q = _noalloc(s, np.zeros((0, 3)), np.zeros(0), q_init=1.0)
        assert np.abs(q).max() == 0.0


def test_noalloc_deposit_rejects_sizes_beyond_its_scratch():
    # 257 x 257 x 65 vertices alone need ~51 MB of bucket counts: refused
    # before anything is launched
    prm = default_rd3d_params(nx=257, ny=257, nz=65, h=2.0, origin=(-256.0, -256.0, 0.0), max_levels=2)
    with Nutrients3D(prm) as s:
        with pytest.raises(BcgError) as e:
            s._check(lib().bcg_rd3d_deposit_noalloc(s._h, ctypes.c_void_p(1), ctypes.c_void_p(1), 10,
                                                    ctypes.c_void_p(1), None))
        assert e.value.code == -2 and "scratch" in str(e.value)
