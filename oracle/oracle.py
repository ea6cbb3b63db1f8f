"""ctypes loaders for the test oracles (TEST INFRASTRUCTURE ONLY).

`Oracle` wraps oracle/liboracle.so, the C restatement of the reference
mechanics (bcg_oracle.c). `RefLib` wraps oracle/_ref/libbcgref.so, the
reference's own sources compiled with a small harness (ref_shim.cpp); it is
only present where the reference checkout was available at build time.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this module, and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from baccolonygrowth_amd.cells import CELL_DTYPE, Params

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbcgref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


def build(ref: bool = True) -> None:
    """Compile the oracle (and the reference harness when its sources exist)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _cells_ptr(cells: np.ndarray):
    assert cells.dtype == CELL_DTYPE and cells.flags.c_contiguous
    return cells.ctypes.data_as(_vp)


def _dptr(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(_dp), a


class _Base:
    prefix = ""

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def min_distance(self, a, b):
        out = np.zeros(7)
        self._f("min_distance")(*self._pre(), _cells_ptr(a), _cells_ptr(b), out.ctypes.data_as(_dp))
        return out

    def F_cc(self, p: Params, a, b):
        out = np.zeros(7)
        self._f("F_cc")(ctypes.byref(p), _cells_ptr(a), _cells_ptr(b), out.ctypes.data_as(_dp))
        return out

    def F_cw(self, p: Params, c, wall_z: float):
        out = np.zeros(18)
        self._f("F_cw")(ctypes.byref(p), _cells_ptr(c), ctypes.c_double(wall_z), out.ctypes.data_as(_dp))
        return out


class Sim:
    """A stateful mechanics simulator from either library (same surface)."""

    def __init__(self, owner, p: Params):
        self.o = owner
        self.p = p.copy()
        self.h = owner._f("sim_create")(ctypes.byref(self.p))

    def close(self):
        if self.h:
            self.o._f("sim_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_cells(self, cells):
        cells = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        rc = self.o._f("sim_set_cells")(self.h, _cells_ptr(cells), len(cells))
        if rc not in (None, 0):
            raise RuntimeError(f"set_cells rc={rc}")

    def set_fields(self, height=None, normal=None, wall=None, gr1=None, gr2=None):
        keep = []
        args = []
        for a in (height, normal, wall, gr1, gr2):
            if a is None:
                args.append(None)
            else:
                ptr, arr = _dptr(a)
                keep.append(arr)
                args.append(ptr)
        self.o._f("sim_set_fields")(self.h, *args)

    @property
    def n(self) -> int:
        return self.o._f("sim_num_cells")(self.h)

    def cells(self):
        out = np.zeros(self.n, CELL_DTYPE)
        self.o._f("sim_get_cells")(self.h, _cells_ptr(out))
        return out

    def rebuild(self, ids=None):
        if ids is None:
            rc = self.o._f("sim_rebuild")(self.h)
        else:
            ids = np.ascontiguousarray(ids, np.int32)
            rc = self.o._f("sim_rebuild_ids")(self.h, ids.ctypes.data_as(_ip), len(ids))
        if rc not in (None, 0):
            raise RuntimeError(f"rebuild rc={rc}")

    def neighbours(self, stride=256):
        n = self.n
        counts = np.zeros(n, np.int32)
        lists = np.full((n, stride), -1, np.int32)
        self.o._f("sim_get_neighbours")(self.h, counts.ctypes.data_as(_ip), lists.ctypes.data_as(_ip), stride)
        return counts, lists

    def boxes(self, stride=256):
        p = self.p
        nbox = p.BoxX * p.BoxY * p.BoxZ
        counts = np.zeros(nbox, np.int32)
        ids = np.full((nbox, stride), -1, np.int32)
        self.o._f("sim_get_boxes")(self.h, counts.ctypes.data_as(_ip), ids.ctypes.data_as(_ip), stride)
        return counts, ids

    def sum_forces(self):
        n = self.n
        F = np.zeros((n, 3))
        T = np.zeros((n, 3))
        self.o._f("sim_sum_forces")(self.h, F.ctypes.data_as(_dp), T.ctypes.data_as(_dp))
        return F, T


    def mean_stress(self):
        """mean_stress of every cell (stress[N, 5], Fnet[N, 3]); takes the
        rand() draws of its F_cw calls."""
        n = self.n
        S = np.zeros((n, 5))
        F = np.zeros((n, 3))
        self.o._f("sim_mean_stress")(self.h, S.ctypes.data_as(_dp), F.ctypes.data_as(_dp))
        return S, F

    def output_cells(self, path: str, t: float) -> None:
        """The reference's own cell file of an output step (reference
        library only)."""
        f = self.o._f("sim_output_cells")
        f.argtypes = [_vp, ctypes.c_char_p, ctypes.c_double]
        if f(self.h, path.encode(), ctypes.c_double(t)) != 0:
            raise OSError(f"cannot write {path}")

    def step_par(self, dt: float, t: float, rebuild: bool, lineage_cap: int = 100000):
        """The OpenMP form of step (C oracle only); the same result."""
        lin = np.zeros((lineage_cap, 2), np.int32)
        nd = self.o.lib.or_sim_step_par(self.h, dt, t, int(bool(rebuild)), lin.ctypes.data_as(_ip), lineage_cap)
        if nd < 0:
            raise RuntimeError(f"step rc={nd}")
        return lin[: min(nd, lineage_cap)].copy()

    def step(self, dt: float, t: float, rebuild: bool, lineage_cap: int = 100000):
        lin = np.zeros((lineage_cap, 2), np.int32)
        nd = self.o._f("sim_step")(
            self.h, ctypes.c_double(dt), ctypes.c_double(t), int(bool(rebuild)), lin.ctypes.data_as(_ip), lineage_cap
        )
        if nd < 0:
            raise RuntimeError(f"step rc={nd}")
        return lin[: min(nd, lineage_cap)].copy()


def _declare(lib, prefix):
    f = lambda n: getattr(lib, prefix + n)  # noqa: E731
    f("sim_create").restype = _vp
    f("sim_create").argtypes = [ctypes.POINTER(Params)]
    f("sim_destroy").argtypes = [_vp]
    f("sim_set_cells").argtypes = [_vp, _vp, ctypes.c_int]
    f("sim_set_fields").argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
    f("sim_num_cells").argtypes = [_vp]
    f("sim_num_cells").restype = ctypes.c_int
    f("sim_get_cells").argtypes = [_vp, _vp]
    f("sim_rebuild").argtypes = [_vp]
    f("sim_rebuild_ids").argtypes = [_vp, _ip, ctypes.c_int]
    f("sim_get_neighbours").argtypes = [_vp, _ip, _ip, ctypes.c_int]
    f("sim_get_boxes").argtypes = [_vp, _ip, _ip, ctypes.c_int]
    f("sim_sum_forces").argtypes = [_vp, _dp, _dp]
    f("sim_mean_stress").argtypes = [_vp, _dp, _dp]
    f("sim_step").argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int]
    f("sim_step").restype = ctypes.c_int
    f("F_cc").argtypes = [ctypes.POINTER(Params), _vp, _vp, _dp]
    f("F_cw").argtypes = [ctypes.POINTER(Params), _vp, ctypes.c_double, _dp]


class Oracle(_Base):
    """The C restatement (liboracle.so)."""

    prefix = "or_"

    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        self.lib = ctypes.CDLL(ORACLE_SO)
        _declare(self.lib, self.prefix)
        self.lib.or_sim_set_cells.restype = ctypes.c_int
        self.lib.or_sim_rebuild.restype = ctypes.c_int
        self.lib.or_sim_rebuild_ids.restype = ctypes.c_int
        self.lib.or_sim_wall_draws.argtypes = [_vp]
        self.lib.or_sim_wall_draws.restype = ctypes.c_int64
        self.lib.or_min_distance.argtypes = [_vp, _vp, _dp]
        self.lib.or_srand.argtypes = [ctypes.c_uint32]
        self.lib.or_rand.restype = ctypes.c_int32
        self.lib.or_sim_step_par.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int]
        self.lib.or_sim_step_par.restype = ctypes.c_int
        self.lib.or_set_threads.argtypes = [ctypes.c_int]
        self.lib.or_get_threads.restype = ctypes.c_int

    def _pre(self):
        return ()

    def sim(self, p: Params) -> Sim:
        return Sim(self, p)

    def srand(self, seed: int):
        """Seed the oracle's own rand() stream (glibc's algorithm, kept apart
        from the process C library so that nothing else can advance it)."""
        self.lib.or_srand(seed)

    def rand(self) -> int:
        return self.lib.or_rand()

    def set_threads(self, n: int) -> None:
        """OpenMP threads of Sim.step_par."""
        self.lib.or_set_threads(n)

    def keep_batch(self, a: np.ndarray, b: np.ndarray, Rc: float) -> np.ndarray:
        """reduceNeighbours' test of the pairs (a[i], b[i]): 1 when b[i] stays
        in a[i]'s list."""
        a = np.ascontiguousarray(a)
        b = np.ascontiguousarray(b)
        assert len(a) == len(b)
        out = np.zeros(len(a), np.int32)
        self.lib.or_keep_batch.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_double, _ip]
        self.lib.or_keep_batch(_cells_ptr(a), _cells_ptr(b), len(a), Rc, out.ctypes.data_as(_ip))
        return out

    def glibc_stream(self, seed: int, n: int) -> np.ndarray:
        class G(ctypes.Structure):
            _fields_ = [("r", ctypes.c_int32 * 34), ("i", ctypes.c_int32)]

        g = G()
        self.lib.or_glibc_srand(ctypes.byref(g), ctypes.c_uint32(seed))
        self.lib.or_glibc_rand.restype = ctypes.c_int32
        return np.array([self.lib.or_glibc_rand(ctypes.byref(g)) for _ in range(n)], np.int64)


class RefLib(_Base):
    """The reference's own code (oracle/_ref/libbcgref.so)."""

    prefix = "ref_"

    def __init__(self):
        if not os.path.exists(REF_SO):
            raise FileNotFoundError(REF_SO)
        self.lib = ctypes.CDLL(REF_SO)
        _declare(self.lib, self.prefix)
        self.lib.ref_min_distance.argtypes = [ctypes.POINTER(Params), _vp, _vp, _dp]
        self.lib.ref_set_params.argtypes = [ctypes.POINTER(Params)]
        self.lib.ref_srand.argtypes = [ctypes.c_uint]
        for name in ("sim_set_cells", "sim_rebuild", "sim_rebuild_ids", "sim_set_fields"):
            getattr(self.lib, "ref_" + name).restype = None
        self.lib.ref_rand.restype = ctypes.c_int
        self.lib.ref_set_threads.argtypes = [ctypes.c_int]
        self.lib.ref_sim_step_omp.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int]
        self.lib.ref_sim_step_omp.restype = ctypes.c_int
        self._p = Params()

    def step_split(self, sim: "Sim", dt: float, t: float, rebuild: bool, lineage_cap: int = 100000):
        """The reference's step that also returns the moved and grown cells
        before the divisions (what its output step saves): (lineage, moved)."""
        L = self.lib
        L.ref_sim_step_split.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _vp, _ip, ctypes.c_int]
        L.ref_sim_step_split.restype = ctypes.c_int
        moved = np.zeros(sim.n, CELL_DTYPE)
        lin = np.zeros((lineage_cap, 2), np.int32)
        nd = L.ref_sim_step_split(sim.h, dt, t, int(bool(rebuild)), _cells_ptr(moved), lin.ctypes.data_as(_ip),
                                  lineage_cap)
        if nd < 0:
            raise RuntimeError(f"step rc={nd}")
        return lin[: min(nd, lineage_cap)].copy(), moved

    def set_threads(self, n: int) -> None:
        """OpenMP threads of the reference's parallel regions (1 for parity)."""
        self.lib.ref_set_threads(n)

    # -- the reference's file writers and readers (ref_shim.cpp, file I/O)
    def write_agar(self, path: str, agar: np.ndarray, BX: int, BY: int, BZA: int, levels: int, n: int = 1) -> None:
        a = np.ascontiguousarray(agar)
        L = self.lib
        L.ref_write_agar.argtypes = [ctypes.c_char_p, _vp] + [ctypes.c_int] * 5
        if L.ref_write_agar(path.encode(), a.ctypes.data, BX, BY, BZA, levels, n):
            raise OSError(path)

    def write_wall(self, path: str, wall: np.ndarray, BX: int, BY: int, levels: int) -> None:
        w = np.ascontiguousarray(wall)
        L = self.lib
        L.ref_write_wall.argtypes = [ctypes.c_char_p, _vp] + [ctypes.c_int] * 3
        if L.ref_write_wall(path.encode(), w.ctypes.data, BX, BY, levels):
            raise OSError(path)

    def read_pair(self, kind: str, path: str, a: np.ndarray, b: np.ndarray, *dims) -> None:
        """Two Input calls on one stream into a then b (kind agar, wall or
        env; dims BX, BY[, BZ]); both arrays are read and written in place."""
        fn = getattr(self.lib, f"ref_read_{kind}_pair")
        fn.argtypes = [ctypes.c_char_p, _vp, _vp] + [ctypes.c_int] * len(dims)
        assert a.flags.c_contiguous and b.flags.c_contiguous
        if fn(path.encode(), a.ctypes.data, b.ctypes.data, *dims):
            raise OSError(path)

    def save_cells(self, path: str, cells: np.ndarray, t: float, dt: float) -> None:
        c = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        L = self.lib
        L.ref_save_cells.argtypes = [ctypes.c_char_p, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        if L.ref_save_cells(path.encode(), c.ctypes.data, len(c), t, dt):
            raise OSError(path)

    def add_first_cells(self, sim: "Sim", size: int = 1, radius: float = 8.0) -> int:
        """AddFirstCells into `sim` (size 1 only)."""
        f = self.lib.ref_sim_add_first_cells
        f.argtypes = [_vp, ctypes.c_int, ctypes.c_double]
        n = f(sim.h, size, radius)
        if n < 0:
            raise ValueError(size)
        return n

    def rough_wall(self, p: Params) -> np.ndarray:
        """The rough wall's expression over this library's rand() stream."""
        w = np.zeros(p.BoxX * p.BoxY)
        self.lib.ref_rough_wall.argtypes = [ctypes.POINTER(Params), _vp]
        self.lib.ref_rough_wall.restype = None
        self.lib.ref_rough_wall(ctypes.byref(p), w.ctypes.data)
        return w

    def sim_rough_wall(self, sim: "Sim") -> None:
        """The run's rough wall into `sim` (rand() draws)."""
        self.lib.ref_sim_rough_wall.argtypes = [_vp]
        self.lib.ref_sim_rough_wall.restype = None
        self.lib.ref_sim_rough_wall(sim.h)

    def load_cells(self, sim: "Sim", path: str):
        """LoadCells into `sim`: (t, dt, number of cells)."""
        L = self.lib
        L.ref_sim_load_cells.argtypes = [_vp, ctypes.c_char_p, _dp, _dp]
        t, dt = ctypes.c_double(0), ctypes.c_double(0)
        n = L.ref_sim_load_cells(sim.h, path.encode(), ctypes.byref(t), ctypes.byref(dt))
        if n < 0:
            raise OSError(path)
        return t.value, dt.value, n

    def step_omp(self, sim: "Sim", dt: float, t: float, rebuild: bool) -> int:
        """The reference's step with its threaded move/grow loop (timing only:
        the division order then depends on the threads)."""
        lin = np.zeros((1, 2), np.int32)
        nd = self.lib.ref_sim_step_omp(sim.h, dt, t, int(bool(rebuild)), lin.ctypes.data_as(_ip), 0)
        if nd < 0:
            raise RuntimeError(f"step rc={nd}")
        return nd

    def _pre(self):
        return (ctypes.byref(self._p),)

    def sim(self, p: Params) -> Sim:
        return Sim(self, p)

    def srand(self, seed: int):
        self.lib.ref_srand(seed)

    def rand(self) -> int:
        return self.lib.ref_rand()


class RefEnv:
    """The reference's field and nutrient passes on a RefLib Sim (test
    infrastructure: ref_env_* of oracle/ref_shim.cpp)."""

    def __init__(self, ref: "RefLib", sim: Sim, env_params):
        from baccolonygrowth_amd.environment import ENV_DTYPE, AGA_DTYPE

        self.ref, self.sim, self.p = ref, sim, env_params
        L = ref.lib
        L.ref_env_create.restype = _vp
        L.ref_env_create.argtypes = [_vp, _vp]
        L.ref_env_destroy.argtypes = [_vp]
        L.ref_env_destroy.restype = None
        L.ref_env_fields.argtypes = [_vp, ctypes.c_int]
        L.ref_env_fields.restype = None
        L.ref_env_nutrients.argtypes = [_vp]
        L.ref_env_nutrients.restype = ctypes.c_int
        L.ref_env_get_fields.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
        L.ref_env_get_fields.restype = None
        L.ref_env_get_state.argtypes = [_vp] * 7
        L.ref_env_get_state.restype = None
        self._env_dtype, self._aga_dtype = ENV_DTYPE, AGA_DTYPE
        self.h = L.ref_env_create(ctypes.byref(env_params), sim.h)

    def close(self):
        if self.h:
            self.ref.lib.ref_env_destroy(self.h)
            self.h = None

    def fields(self, pass_: int):
        self.ref.lib.ref_env_fields(self.h, pass_)

    def output(self, density_path: str, env_path: str) -> None:
        """The reference's own density and environment files."""
        f = self.ref.lib.ref_env_output
        f.argtypes = [_vp, ctypes.c_char_p, ctypes.c_char_p]
        f.restype = ctypes.c_int
        if f(self.h, density_path.encode(), env_path.encode()) != 0:
            raise OSError("cannot write the output files")

    def nutrients(self) -> int:
        return self.ref.lib.ref_env_nutrients(self.h)

    def append(self, agar_path: str, wall_path: str) -> None:
        """The reference's own AgarField and WallField files of the state."""
        f = self.ref.lib.ref_env_append
        f.argtypes = [_vp, ctypes.c_char_p, ctypes.c_char_p]
        if f(self.h, agar_path.encode(), wall_path.encode()) != 0:
            raise OSError("cannot write the field files")

    def field_restart(self, dirname: str, index: int) -> None:
        """The reference's field restart reads from <dirname>/.../<index>.txt."""
        f = self.ref.lib.ref_env_field_restart
        f.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int]
        if f(self.h, dirname.encode(), index) != 0:
            raise OSError(dirname)

    def get_fields(self):
        p = self.p
        nb = p.BoxX * p.BoxZ
        nh = p.BoxX * p.refinementGridHeight * p.BoxY * p.refinementGridHeight
        out = {k: np.zeros(nb) for k in ("density", "density_shift", "inside")}
        out["wall_shift"] = np.zeros(p.BoxX)
        out["height"] = np.zeros(nh)
        out["normal"] = np.zeros(3 * nh)
        b = np.zeros(5, dtype=np.int32)
        P = lambda a: a.ctypes.data_as(_dp)  # noqa: E731
        self.ref.lib.ref_env_get_fields(self.h, P(out["density"]), P(out["density_shift"]), P(out["inside"]),
                                        P(out["wall_shift"]), P(out["height"]), P(out["normal"]),
                                        b.ctypes.data_as(_ip))
        out["bounds"] = b
        return out

    def get_state(self):
        p = self.p
        nb, na = p.BoxX * p.BoxZ, p.BoxX * p.BoxZAgar
        st = {
            "colony": np.zeros(nb, self._env_dtype),
            "colony_old": np.zeros(nb, self._env_dtype),
            "agar": np.zeros(p.maxLevelsMG * na, self._aga_dtype),
            "agar_old": np.zeros(p.maxLevelsMG * na, self._aga_dtype),
            "wall": np.zeros(p.maxLevelsMG * p.BoxX, self._aga_dtype),
            "wall_old": np.zeros(p.maxLevelsMG * p.BoxX, self._aga_dtype),
        }
        self.ref.lib.ref_env_get_state(self.h, *[ctypes.c_void_p(a.ctypes.data) for a in st.values()])
        return st


def ref_available() -> bool:
    return os.path.exists(REF_SO)
