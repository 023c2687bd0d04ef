"""MI355X-native multispin +-J Edwards-Anderson Metropolis engine.

The compute path is the in-tree HIP library ``libsa_hip.so`` (kernels for gfx950,
the C ABI of ``include/sa_hip.h`` and the C++ drivers of ``include/sa``).  This
module is a thin ctypes binding of that C ABI for Python callers (tests, bench).
There is no fallback: when the library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsa_hip.so")

KIND_2P, KIND_2D, KIND_3D = 0, 1, 2
RNG_REPLAY, RNG_COUNTER = 0, 1
SA_OK = 0
TUNE_SYS_PER_THREAD, TUNE_SWEEP3D_KERNEL, TUNE_ROWS_PER_TILE, TUNE_XCD_MAP = 1, 2, 3, 4
TUNE_BATCH_BLOCK, TUNE_BATCH_CHUNK = 101, 102

# every symbol include/sa_hip.h declares (checked by tests/test_capi_symbols.py)
EXPORTS = [
    "sa_last_error", "sa_status_string", "sa_version",
    "sa_ladder_create", "sa_ladder_destroy", "sa_ladder_sites", "sa_ladder_nsys",
    "sa_ladder_set_rng", "sa_ladder_seed", "sa_ladder_get_sweep_index", "sa_ladder_set_sweep_index",
    "sa_ladder_get_engine", "sa_ladder_set_engine",
    "sa_ladder_init_couplings", "sa_ladder_init_spins", "sa_ladder_set_couplings", "sa_ladder_get_couplings",
    "sa_ladder_set_spins", "sa_ladder_get_spins", "sa_ladder_set_temperature", "sa_ladder_sweep",
    "sa_ladder_anneal",
    "sa_ladder_measure", "sa_ladder_measure_systems", "sa_ladder_swap", "sa_ladder_sync", "sa_ladder_stream",
    "sa_ladder_set_tuning", "sa_ladder_get_tuning",
    "sa_pt_run", "sa_anneal_run",
    "sa_pt_create", "sa_pt_step", "sa_pt_sync", "sa_pt_steps_done", "sa_pt_result", "sa_pt_ladder",
    "sa_pt_destroy",
    "sa_batch_create", "sa_batch_destroy", "sa_batch_add", "sa_batch_run", "sa_batch_sweeps_done",
    "sa_batch_get_spins", "sa_batch_measure", "sa_batch_ladder", "sa_batch_remove", "sa_batch_stream",
    "sa_batch_set_tuning", "sa_batch_get_tuning",
]


class SAError(RuntimeError):
    pass


_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_lib = None


def lib():
    """Load libsa_hip.so (built by __graft_entry__.build()); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SAError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(python -m simulated_annealing_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.sa_last_error.restype = C.c_char_p
        L.sa_status_string.restype = C.c_char_p
        L.sa_status_string.argtypes = [C.c_int]
        L.sa_ladder_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]
        L.sa_ladder_destroy.argtypes = [_vp]
        L.sa_ladder_sites.argtypes = [_vp]
        L.sa_ladder_sites.restype = C.c_int64
        L.sa_ladder_nsys.argtypes = [_vp]
        L.sa_ladder_set_rng.argtypes = [_vp, C.c_int]
        L.sa_ladder_seed.argtypes = [_vp, C.c_uint64]
        L.sa_ladder_get_sweep_index.argtypes = [_vp, C.POINTER(C.c_uint64)]
        L.sa_ladder_set_sweep_index.argtypes = [_vp, C.c_uint64]
        L.sa_ladder_get_engine.argtypes = [_vp, _u64p]
        L.sa_ladder_set_engine.argtypes = [_vp, _u64p]
        L.sa_ladder_init_couplings.argtypes = [_vp]
        L.sa_ladder_init_spins.argtypes = [_vp]
        L.sa_ladder_set_couplings.argtypes = [_vp, _u64p, _u64p, _vp]
        L.sa_ladder_get_couplings.argtypes = [_vp, _u64p, _u64p, _vp]
        L.sa_ladder_set_spins.argtypes = [_vp, C.c_int, _u64p]
        L.sa_ladder_get_spins.argtypes = [_vp, C.c_int, _u64p]
        L.sa_ladder_set_temperature.argtypes = [_vp, C.c_int, C.c_double, C.c_double, _vp]
        L.sa_ladder_sweep.argtypes = [_vp, C.c_int]
        L.sa_ladder_anneal.argtypes = [_vp, _f64p, C.c_int, C.c_double]
        L.sa_ladder_measure.argtypes = [_vp, _i32p, _i32p]
        L.sa_ladder_measure_systems.argtypes = [_vp, _i32p, C.c_int, _i32p, _i32p]
        L.sa_ladder_swap.argtypes = [_vp, C.c_int, C.c_int, C.c_uint64]
        L.sa_ladder_sync.argtypes = [_vp]
        L.sa_ladder_stream.argtypes = [_vp]
        L.sa_ladder_stream.restype = _vp
        L.sa_ladder_set_tuning.argtypes = [_vp, C.c_int, C.c_int]
        L.sa_ladder_get_tuning.argtypes = [_vp, C.c_int, C.POINTER(C.c_int)]
        L.sa_pt_run.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                _vp, _vp, _vp, _vp, _vp]
        L.sa_anneal_run.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int,
                                    _vp, _vp]
        L.sa_pt_create.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_uint64, C.c_int, C.c_int,
                                   C.POINTER(_vp)]
        L.sa_pt_step.argtypes = [_vp, C.c_int]
        L.sa_pt_sync.argtypes = [_vp]
        L.sa_pt_steps_done.argtypes = [_vp]
        L.sa_pt_steps_done.restype = C.c_int64
        L.sa_pt_result.argtypes = [_vp, C.c_int64, _vp, _vp, _vp, _vp, _vp]
        L.sa_pt_ladder.argtypes = [_vp]
        L.sa_pt_ladder.restype = _vp
        L.sa_pt_destroy.argtypes = [_vp]
        L.sa_batch_create.argtypes = [C.POINTER(_vp)]
        L.sa_batch_destroy.argtypes = [_vp]
        L.sa_batch_add.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, _f64p, C.c_int, C.c_double,
                                   C.POINTER(C.c_int)]
        L.sa_batch_run.argtypes = [_vp, C.c_int]
        L.sa_batch_sweeps_done.argtypes = [_vp, C.c_int, C.POINTER(C.c_int)]
        L.sa_batch_get_spins.argtypes = [_vp, C.c_int, C.c_int, _u64p]
        L.sa_batch_measure.argtypes = [_vp, C.c_int, _i32p, _i32p]
        L.sa_batch_ladder.argtypes = [_vp, C.c_int]
        L.sa_batch_ladder.restype = _vp
        L.sa_batch_remove.argtypes = [_vp, C.c_int]
        L.sa_batch_stream.argtypes = [_vp]
        L.sa_batch_stream.restype = _vp
        L.sa_batch_set_tuning.argtypes = [_vp, C.c_int, C.c_int]
        L.sa_batch_get_tuning.argtypes = [_vp, C.c_int, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        L = lib()
        raise SAError(f"{what}: {L.sa_status_string(rc).decode()} ({L.sa_last_error().decode()})")
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


def nsites(kind: int, L: int) -> int:
    return L * L * (L if kind == KIND_3D else 1)


class Ladder:
    """`nsys` systems of one geometry in HBM sharing couplings and random stream."""

    def __init__(self, kind: int, L: int, nsys: int = 1):
        self.kind, self.L, self.nsys = kind, L, nsys
        self.N = nsites(kind, L)
        h = _vp()
        check(lib().sa_ladder_create(kind, L, nsys, C.byref(h)), "sa_ladder_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().sa_ladder_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- streams / init
    def set_rng(self, mode: int):
        check(lib().sa_ladder_set_rng(self.h, mode), "set_rng")

    def seed(self, seed: int):
        check(lib().sa_ladder_seed(self.h, seed), "seed")

    def seed_like_main(self, seed: int):
        """set_seed, init_J, set_seed, init_rand (src/main_gpu.cu:262-276)."""
        self.seed(seed)
        self.init_couplings()
        self.seed(seed)
        self.init_spins()

    def init_couplings(self):
        check(lib().sa_ladder_init_couplings(self.h), "init_couplings")

    def init_spins(self):
        check(lib().sa_ladder_init_spins(self.h), "init_spins")

    @property
    def sweep_index(self) -> int:
        t = C.c_uint64()
        check(lib().sa_ladder_get_sweep_index(self.h, C.byref(t)), "sweep_index")
        return t.value

    @sweep_index.setter
    def sweep_index(self, t: int):
        check(lib().sa_ladder_set_sweep_index(self.h, t), "sweep_index")

    def engine(self) -> np.ndarray:
        st = np.zeros(313, dtype=np.uint64)
        check(lib().sa_ladder_get_engine(self.h, st), "get_engine")
        return st

    def set_engine(self, st):
        check(lib().sa_ladder_set_engine(self.h, np.ascontiguousarray(st, dtype=np.uint64)), "set_engine")

    # ---- data
    def set_couplings(self, jx, jy, jz=None):
        jz_arr = None if jz is None else np.ascontiguousarray(jz, dtype=np.uint64)
        check(lib().sa_ladder_set_couplings(self.h, np.ascontiguousarray(jx, dtype=np.uint64),
                                            np.ascontiguousarray(jy, dtype=np.uint64), _ptr(jz_arr)),
              "set_couplings")

    def couplings(self):
        jx = np.zeros(self.N, dtype=np.uint64)
        jy = np.zeros(self.N, dtype=np.uint64)
        jz = np.zeros(self.N, dtype=np.uint64)
        check(lib().sa_ladder_get_couplings(self.h, jx, jy, _ptr(jz)), "get_couplings")
        return jx, jy, jz

    def set_spins(self, sys: int, s):
        check(lib().sa_ladder_set_spins(self.h, sys, np.ascontiguousarray(s, dtype=np.uint64)), "set_spins")

    def spins(self, sys: int = 0) -> np.ndarray:
        s = np.zeros(self.N, dtype=np.uint64)
        check(lib().sa_ladder_get_spins(self.h, sys, s), "get_spins")
        return s

    def set_temperature(self, sys: int, T: float, h: float = 0.0) -> np.ndarray:
        b = np.zeros(20, dtype=np.float64)
        check(lib().sa_ladder_set_temperature(self.h, sys, T, h, _ptr(b)), "set_temperature")
        return b[: 14 if self.kind == KIND_3D else 20]

    # ---- work
    def sweep(self, n: int = 1):
        check(lib().sa_ladder_sweep(self.h, n), "sweep")

    def anneal(self, schedule, h: float = 0.0):
        """One sweep per row of `schedule` ([nsweeps] or [nsweeps, nsys]) at those temperatures."""
        T = np.asarray(schedule, dtype=np.float64)
        if T.ndim == 1:
            T = np.repeat(T[:, None], self.nsys, axis=1)
        T = np.ascontiguousarray(T.reshape(-1, self.nsys))
        check(lib().sa_ladder_anneal(self.h, T.reshape(-1) if T.size else np.zeros(1), T.shape[0], h), "anneal")

    def measure(self):
        """(E, M) int32 arrays [nsys, 64] in lane order."""
        E = np.zeros(self.nsys * 64, dtype=np.int32)
        M = np.zeros(self.nsys * 64, dtype=np.int32)
        check(lib().sa_ladder_measure(self.h, E, M), "measure")
        return E.reshape(self.nsys, 64), M.reshape(self.nsys, 64)

    def measure_systems(self, systems):
        """(E, M) [len(systems), 64] for the listed systems only."""
        idx = np.ascontiguousarray(systems, dtype=np.int32)
        E = np.zeros(idx.size * 64, dtype=np.int32)
        M = np.zeros(idx.size * 64, dtype=np.int32)
        check(lib().sa_ladder_measure_systems(self.h, idx, idx.size, E, M), "measure_systems")
        return E.reshape(-1, 64), M.reshape(-1, 64)

    def swap(self, a: int, b: int, mask: int):
        check(lib().sa_ladder_swap(self.h, a, b, C.c_uint64(mask & 0xFFFFFFFFFFFFFFFF)), "swap")

    def sync(self):
        check(lib().sa_ladder_sync(self.h), "sync")

    def set_sys_per_thread(self, mb: int):
        self.tune(TUNE_SYS_PER_THREAD, mb)

    def sys_per_thread(self) -> int:
        return self.tuning(TUNE_SYS_PER_THREAD)

    def tune(self, key: int, value: int):
        check(lib().sa_ladder_set_tuning(self.h, key, value), "set_tuning")

    def tuning(self, key: int) -> int:
        v = C.c_int()
        check(lib().sa_ladder_get_tuning(self.h, key, C.byref(v)), "get_tuning")
        return v.value


class PT:
    """A parallel-tempering run on one ladder (sa_pt_* of include/sa_hip.h)."""

    def __init__(self, kind, L, T, h=0.0, seed=1234, sweeps=10, rng_mode=RNG_COUNTER):
        self.kind, self.L = kind, L
        self.T = np.ascontiguousarray(T, dtype=np.float64)
        self.nT = self.T.size
        self.N = nsites(kind, L)
        p = _vp()
        check(lib().sa_pt_create(kind, L, self.nT, self.T, h, seed, sweeps, rng_mode, C.byref(p)), "sa_pt_create")
        self.p = p

    def step(self, n=1):
        check(lib().sa_pt_step(self.p, n), "sa_pt_step")

    def sync(self):
        check(lib().sa_pt_sync(self.p), "sa_pt_sync")

    def ladder_handle(self):
        return lib().sa_pt_ladder(self.p)

    def result(self, spins=True):
        n = lib().sa_pt_steps_done(self.p)
        pairs = np.zeros(max(n, 1), dtype=np.int32)
        us = np.zeros(max(n, 1), dtype=np.float64)
        masks = np.zeros(max(n, 1), dtype=np.uint64)
        E = np.zeros(self.nT * 64, dtype=np.float32)
        sp = np.zeros(self.nT * self.N, dtype=np.uint64) if spins else None
        check(lib().sa_pt_result(self.p, n, _ptr(pairs), _ptr(us), _ptr(masks), _ptr(E), _ptr(sp)), "sa_pt_result")
        out = dict(pairs=pairs[:n], us=us[:n], masks=masks[:n], E=E.reshape(self.nT, 64))
        if spins:
            out["spins"] = sp.reshape(self.nT, self.N)
        return out

    def close(self):
        if getattr(self, "p", None):
            lib().sa_pt_destroy(self.p)
            self.p = None

    def __del__(self):
        self.close()


def linear_schedule(T0: float, T1: float, nsweeps: int) -> np.ndarray:
    """Temperature of sweep t of a linear anneal from T0 to T1 (t = 0 .. nsweeps-1)."""
    if nsweeps <= 1:
        return np.full(max(nsweeps, 0), T0, dtype=np.float64)
    t = np.arange(nsweeps, dtype=np.float64)
    return T0 + (T1 - T0) * t / (nsweeps - 1)


class Batch:
    """Independent instances of uneven size swept together (sa_batch_* of include/sa_hip.h)."""

    def __init__(self):
        b = _vp()
        check(lib().sa_batch_create(C.byref(b)), "sa_batch_create")
        self.b = b
        self.info = {}

    def add(self, kind, L, seed, nsweeps, T=None, schedule=None, nsys=1, h=0.0) -> int:
        """Add an instance of `nsys` systems seeded like main (set_seed, init_J,
        set_seed, init_rand).  Give either T: [nsys] fixed temperatures, or
        schedule: [nsweeps] or [nsweeps, nsys], the temperatures of every sweep."""
        if (T is None) == (schedule is None):
            raise ValueError("give exactly one of T and schedule")
        per_sweep = schedule is not None
        arr = np.asarray(schedule if per_sweep else T, dtype=np.float64)
        if per_sweep and arr.ndim == 1:
            arr = np.repeat(arr[:, None], nsys, axis=1)
        need = nsweeps * nsys if per_sweep else nsys
        Tc = np.ascontiguousarray(arr.reshape(-1))
        if Tc.size != need:
            raise ValueError(f"expected {need} temperatures, got {Tc.size}")
        if Tc.size == 0:
            Tc = np.zeros(1)
        i = C.c_int()
        check(lib().sa_batch_add(self.b, kind, L, nsys, seed, nsweeps, Tc, int(per_sweep), h, C.byref(i)),
              "sa_batch_add")
        self.info[i.value] = (kind, L, nsys)
        return i.value

    def run(self, sweeps: int = 0) -> int:
        return check(lib().sa_batch_run(self.b, sweeps), "sa_batch_run")

    def sweeps_done(self, i: int) -> int:
        t = C.c_int()
        check(lib().sa_batch_sweeps_done(self.b, i, C.byref(t)), "sa_batch_sweeps_done")
        return t.value

    def spins(self, i: int, sys: int = 0) -> np.ndarray:
        kind, L, _ = self.info[i]
        s = np.zeros(nsites(kind, L), dtype=np.uint64)
        check(lib().sa_batch_get_spins(self.b, i, sys, s), "sa_batch_get_spins")
        return s

    def measure(self, i: int):
        nsys = self.info[i][2]
        E = np.zeros(nsys * 64, dtype=np.int32)
        M = np.zeros(nsys * 64, dtype=np.int32)
        check(lib().sa_batch_measure(self.b, i, E, M), "sa_batch_measure")
        return E.reshape(nsys, 64), M.reshape(nsys, 64)

    def remove(self, i: int):
        check(lib().sa_batch_remove(self.b, i), "sa_batch_remove")
        self.info.pop(i, None)

    def stream(self):
        return lib().sa_batch_stream(self.b)

    def tune(self, key: int, value: int):
        check(lib().sa_batch_set_tuning(self.b, key, value), "sa_batch_set_tuning")

    def tuning(self, key: int) -> int:
        v = C.c_int()
        check(lib().sa_batch_get_tuning(self.b, key, C.byref(v)), "sa_batch_get_tuning")
        return v.value

    def close(self):
        if getattr(self, "b", None):
            lib().sa_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        self.close()


def pt_run(kind, L, T, h, seed, sweeps, nsteps, rng_mode=RNG_REPLAY):
    T = np.ascontiguousarray(T, dtype=np.float64)
    nT = T.size
    N = nsites(kind, L)
    pairs = np.zeros(nsteps, dtype=np.int32)
    us = np.zeros(nsteps, dtype=np.float64)
    masks = np.zeros(nsteps, dtype=np.uint64)
    E = np.zeros(nT * 64, dtype=np.float32)
    spins = np.zeros(nT * N, dtype=np.uint64)
    check(lib().sa_pt_run(kind, L, nT, T, h, seed, sweeps, nsteps, rng_mode, _ptr(pairs), _ptr(us), _ptr(masks),
                          _ptr(E), _ptr(spins)), "sa_pt_run")
    return dict(pairs=pairs, us=us, masks=masks, E=E.reshape(nT, 64), spins=spins.reshape(nT, N))


def anneal_run(kind, L, Ts, h, seed, stride, rng_mode=RNG_REPLAY):
    Ts = np.ascontiguousarray(Ts, dtype=np.float64)
    nrec = (Ts.size + stride - 1) // stride
    E = np.zeros(max(nrec, 1) * 64, dtype=np.float32)
    spins = np.zeros(nsites(kind, L), dtype=np.uint64)
    n = check(lib().sa_anneal_run(kind, L, Ts, Ts.size, h, seed, stride, rng_mode, _ptr(E), _ptr(spins)),
              "sa_anneal_run")
    return E.reshape(-1, 64)[:n], spins
