"""CPU-side checks of the C-ABI library: it loads, exports every entry point
that include/*.h declares, and its host-only pieces (parameters, the rand()
stream with jump-ahead) behave. No device calls."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from baccolonygrowth_amd import _lib
from baccolonygrowth_amd.cells import read_parameters
from helpers import DEMO_INPUT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        names |= set(re.findall(r"^\s*int\s+(bcg_\w+)\s*\(", txt, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    decl = declared_symbols()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(L, name), name
    # and the Python binding knows every one of them
    assert set(decl) <= set(_lib.exported_symbols())


def test_default_params_match_demo_input():
    from baccolonygrowth_amd.simulation import default_params

    p = default_params()
    q, _ = read_parameters(DEMO_INPUT)
    for k, _t in p._fields_:
        if k in ("maxNeighbours", "planar"):
            continue
        assert getattr(p, k) == getattr(q, k), k


def _stream(seed, skip, n):
    out = np.zeros(n, np.int32)
    rc = _lib.lib().bcg_rand_stream(seed, skip, n, out.ctypes.data_as(_lib._ip))
    assert rc == 0
    return out


def test_rand_stream_matches_host_libc():
    libc = ctypes.CDLL(None)
    for seed in (1, 7, 4242):
        libc.srand(seed)
        host = np.array([libc.rand() for _ in range(2000)], np.int32)
        assert np.array_equal(_stream(seed, 0, 2000), host)


def test_rand_stream_in_bulk_matches_host_libc():
    # 20000 draws: the middle third comes from the bulk generator in several
    # 4096-draw chunks, the last third one at a time after it
    libc = ctypes.CDLL(None)
    libc.srand(5)
    host = np.array([libc.rand() for _ in range(20000)], np.int32)
    assert np.array_equal(_stream(5, 0, 20000), host)


@pytest.mark.parametrize("skip", [1023, 1024, 1025, 5000, 123457])
def test_rand_jump_ahead_equals_stepping(skip):
    libc = ctypes.CDLL(None)
    libc.srand(99)
    for _ in range(skip):
        libc.rand()
    host = np.array([libc.rand() for _ in range(100)], np.int32)
    assert np.array_equal(_stream(99, skip, 100), host)


def test_rand_jump_ahead_large_consistent():
    # jumping 10^7 + k equals jumping 10^7 then reading k more
    a = _stream(3, 10_000_000, 64)
    b = _stream(3, 10_000_032, 32)
    assert np.array_equal(a[32:], b)
