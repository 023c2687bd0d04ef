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
TUNE_BATCH_BLOCK, TUNE_BATCH_CHUNK, TUNE_BATCH_FUSED = 101, 102, 103

# every symbol include/sa_hip.h declares (checked by tests/test_capi_symbols.py)
EXPORTS = [
    "sa_last_error", "sa_status_string", "sa_version",
    "sa_ladder_create", "sa_ladder_destroy", "sa_ladder_sites", "sa_ladder_nsys",
    "sa_ladder_set_rng", "sa_ladder_seed", "sa_ladder_get_seed", "sa_ladder_get_sweep_index",
    "sa_ladder_set_sweep_index",
    "sa_ladder_get_engine", "sa_ladder_set_engine",
    "sa_ladder_init_couplings", "sa_ladder_init_spins", "sa_ladder_set_couplings", "sa_ladder_get_couplings",
    "sa_ladder_load_J", "sa_ladder_save_J",
    "sa_ladder_set_spins", "sa_ladder_get_spins", "sa_ladder_set_temperature", "sa_ladder_sweep",
    "sa_ladder_anneal",
    "sa_ladder_measure", "sa_ladder_measure_systems", "sa_measure_noalloc", "sa_ladder_swap", "sa_ladder_sync", "sa_ladder_stream",
    "sa_ladder_set_tuning", "sa_ladder_get_tuning",
    "sa_pt_run", "sa_anneal_run", "sa_anneal_geometric",
    "sa_pt_create", "sa_pt_step", "sa_pt_sync", "sa_pt_steps_done", "sa_pt_result", "sa_pt_ladder",
    "sa_pt_save", "sa_pt_load", "sa_pt_destroy", "sa_pt_set_output", "sa_pt_step_stride", "sa_pt_print_min",
    "sa_pt_print_swap", "sa_pt_close_output", "sa_pt_update_temperatures", "sa_pt_reset", "sa_pt_temperatures",
    "sa_pt_acceptance", "sa_pt_E_min",
    "sa_batch_create", "sa_batch_destroy", "sa_batch_add", "sa_batch_add_pt", "sa_batch_run",
    "sa_batch_sweeps_done", "sa_batch_get_spins", "sa_batch_measure", "sa_batch_pt_result", "sa_batch_bytes",
    "sa_batch_ladder", "sa_batch_remove", "sa_batch_stream", "sa_batch_set_tuning", "sa_batch_get_tuning",
    "sa_serve_create", "sa_serve_destroy", "sa_serve_clock", "sa_serve_submit", "sa_serve_run", "sa_serve_results",
    "sa_serve_get_stats", "sa_request_temperatures", "sa_request_bytes", "sa_digest",
]

REQ_ANNEAL, REQ_PT = 0, 1


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("L", C.c_int), ("mode", C.c_int), ("nT", C.c_int), ("nsweeps", C.c_int),
                ("every", C.c_int), ("T0", C.c_double), ("T1", C.c_double), ("h", C.c_double),
                ("seed", C.c_uint64), ("arrival", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("steps", C.c_int), ("arrival", C.c_double),
                ("admitted", C.c_double), ("finished", C.c_double), ("bytes", C.c_int64),
                ("attempts", C.c_double), ("spins_digest", C.c_uint64), ("log_digest", C.c_uint64),
                ("E", C.c_int32 * 64), ("M", C.c_int32 * 64), ("E_min", C.c_float * 64)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_active", C.c_int), ("chunk", C.c_int), ("digests", C.c_int),
                ("prep_threads", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("finished", C.c_int64), ("pending", C.c_int64), ("active", C.c_int),
                ("peak_active", C.c_int), ("bytes", C.c_int64), ("peak_bytes", C.c_int64),
                ("hbm_budget", C.c_int64), ("busy_s", C.c_double), ("clock_s", C.c_double),
                ("admit_s", C.c_double), ("retire_s", C.c_double), ("chunks", C.c_int64)]


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
        L.sa_ladder_get_seed.argtypes = [_vp, C.POINTER(C.c_uint64)]
        L.sa_ladder_get_sweep_index.argtypes = [_vp, C.POINTER(C.c_uint64)]
        L.sa_ladder_set_sweep_index.argtypes = [_vp, C.c_uint64]
        L.sa_ladder_get_engine.argtypes = [_vp, _u64p]
        L.sa_ladder_set_engine.argtypes = [_vp, _u64p]
        L.sa_ladder_init_couplings.argtypes = [_vp]
        L.sa_ladder_init_spins.argtypes = [_vp]
        L.sa_ladder_set_couplings.argtypes = [_vp, _u64p, _u64p, _vp]
        L.sa_ladder_get_couplings.argtypes = [_vp, _u64p, _u64p, _vp]
        L.sa_ladder_load_J.argtypes = [_vp, C.c_char_p]
        L.sa_ladder_save_J.argtypes = [_vp, C.c_char_p]
        L.sa_ladder_set_spins.argtypes = [_vp, C.c_int, _u64p]
        L.sa_ladder_get_spins.argtypes = [_vp, C.c_int, _u64p]
        L.sa_ladder_set_temperature.argtypes = [_vp, C.c_int, C.c_double, C.c_double, _vp]
        L.sa_ladder_sweep.argtypes = [_vp, C.c_int]
        L.sa_ladder_anneal.argtypes = [_vp, _f64p, C.c_int, C.c_double]
        L.sa_ladder_measure.argtypes = [_vp, _i32p, _i32p]
        L.sa_ladder_measure_systems.argtypes = [_vp, _i32p, C.c_int, _i32p, _i32p]
        L.sa_measure_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
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
        L.sa_anneal_geometric.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                          C.c_double, C.c_uint64, C.c_int, C.c_int, C.c_char_p, C.c_int64]
        L.sa_anneal_geometric.restype = C.c_int64
        L.sa_pt_create.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_uint64, C.c_int, C.c_int,
                                   C.POINTER(_vp)]
        L.sa_pt_step.argtypes = [_vp, C.c_int]
        L.sa_pt_sync.argtypes = [_vp]
        L.sa_pt_steps_done.argtypes = [_vp]
        L.sa_pt_steps_done.restype = C.c_int64
        L.sa_pt_result.argtypes = [_vp, C.c_int64, _vp, _vp, _vp, _vp, _vp]
        L.sa_pt_ladder.argtypes = [_vp]
        L.sa_pt_ladder.restype = _vp
        L.sa_pt_set_output.argtypes = [_vp, C.c_char_p]
        L.sa_pt_step_stride.argtypes = [_vp, C.c_int, C.c_int]
        L.sa_pt_print_min.argtypes = [_vp]
        L.sa_pt_print_swap.argtypes = [_vp, C.c_char_p, C.c_int64]
        L.sa_pt_print_swap.restype = C.c_int64
        L.sa_pt_close_output.argtypes = [_vp]
        L.sa_pt_save.argtypes = [_vp, C.c_char_p]
        L.sa_pt_load.argtypes = [_vp, C.c_char_p]
        L.sa_pt_destroy.argtypes = [_vp]
        L.sa_pt_update_temperatures.argtypes = [_vp]
        L.sa_pt_reset.argtypes = [_vp]
        L.sa_pt_temperatures.argtypes = [_vp, C.c_int, _vp]
        L.sa_pt_acceptance.argtypes = [_vp, C.c_int, _vp]
        L.sa_pt_E_min.argtypes = [_vp, _vp]
        L.sa_batch_create.argtypes = [C.POINTER(_vp)]
        L.sa_batch_destroy.argtypes = [_vp]
        L.sa_batch_add.argtypes = [_vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, _f64p, C.c_int, C.c_double,
                                   C.POINTER(C.c_int)]
        L.sa_batch_add_pt.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_uint64, C.c_int,
                                      C.c_int, C.POINTER(C.c_int)]
        L.sa_batch_pt_result.argtypes = [_vp, C.c_int, C.c_int64, _vp, _vp, _vp, _vp, _vp]
        L.sa_batch_bytes.argtypes = [_vp]
        L.sa_batch_bytes.restype = C.c_int64
        L.sa_batch_run.argtypes = [_vp, C.c_int]
        L.sa_batch_sweeps_done.argtypes = [_vp, C.c_int, C.POINTER(C.c_int)]
        L.sa_batch_get_spins.argtypes = [_vp, C.c_int, C.c_int, _u64p]
        L.sa_batch_measure.argtypes = [_vp, C.c_int, _i32p, _i32p]
        L.sa_batch_ladder.argtypes = [_vp, C.c_int]
        L.sa_batch_ladder.restype = _vp
        L.sa_batch_remove.argtypes = [_vp, C.c_int]
        L.sa_batch_stream.argtypes = [_vp]
        L.sa_batch_stream.restype = _vp
        L.sa_serve_create.argtypes = [C.POINTER(ServeConfig), C.POINTER(_vp)]
        L.sa_serve_destroy.argtypes = [_vp]
        L.sa_serve_clock.argtypes = [_vp]
        L.sa_serve_clock.restype = C.c_double
        L.sa_serve_submit.argtypes = [_vp, C.POINTER(Request), C.POINTER(C.c_int64)]
        L.sa_serve_run.argtypes = [_vp, C.c_int64]
        L.sa_serve_results.argtypes = [_vp, C.c_int64, C.c_int64, C.POINTER(RequestResult)]
        L.sa_serve_results.restype = C.c_int64
        L.sa_serve_get_stats.argtypes = [_vp, C.POINTER(ServeStats)]
        L.sa_request_temperatures.argtypes = [C.POINTER(Request), _vp, C.c_int64]
        L.sa_request_bytes.argtypes = [C.POINTER(Request)]
        L.sa_request_bytes.restype = C.c_int64
        L.sa_digest.argtypes = [_u64p, C.c_int64, C.c_uint64]
        L.sa_digest.restype = C.c_uint64
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

    def load_J(self, path: str):
        check(lib().sa_ladder_load_J(self.h, os.fsencode(path)), "load_J")

    def save_J(self, path: str):
        check(lib().sa_ladder_save_J(self.h, os.fsencode(path)), "save_J")

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

    def set_output(self, pattern):
        """File-name pattern of the energy files (% -> T<temperature>, E_min)."""
        check(lib().sa_pt_set_output(self.p, str(pattern).encode()), "sa_pt_set_output")

    def step_stride(self, n, stride):
        """n x parallel_tempering::step(stride): energies printed every `stride` sweeps."""
        check(lib().sa_pt_step_stride(self.p, n, stride), "sa_pt_step_stride")

    def print_min(self):
        check(lib().sa_pt_print_min(self.p), "sa_pt_print_min")

    def swap_report(self) -> str:
        n = check(lib().sa_pt_print_swap(self.p, None, 0), "sa_pt_print_swap")
        buf = C.create_string_buffer(n + 1)
        lib().sa_pt_print_swap(self.p, buf, n + 1)
        return buf.value.decode()

    def close_output(self):
        check(lib().sa_pt_close_output(self.p), "sa_pt_close_output")

    def save(self, path):
        """parallel_tempering::save to a file (.tempsys)."""
        check(lib().sa_pt_save(self.p, str(path).encode()), "sa_pt_save")

    def load(self, path):
        """parallel_tempering::load from a file (.tempsys); clears the swap log."""
        check(lib().sa_pt_load(self.p, str(path).encode()), "sa_pt_load")

    def update_temperatures(self) -> int:
        """parallel_tempering::update_T: insert temperatures where the exchange rate
        is below 0.3; returns the new temperature count."""
        n = lib().sa_pt_update_temperatures(self.p)
        check(min(n, 0), "sa_pt_update_temperatures")
        self.nT = n
        return n

    def reset(self):
        """parallel_tempering::reset: E_min cleared, new spins drawn."""
        check(lib().sa_pt_reset(self.p), "sa_pt_reset")

    def temperatures(self) -> np.ndarray:
        out = np.zeros(64, dtype=np.float64)
        n = lib().sa_pt_temperatures(self.p, out.size, _ptr(out))
        check(min(n, 0), "sa_pt_temperatures")
        if n > out.size:
            out = np.zeros(n, dtype=np.float64)
            lib().sa_pt_temperatures(self.p, out.size, _ptr(out))
        return out[:n].copy()

    def acceptance(self) -> np.ndarray:
        """exchange rate of each neighbour pair (num_switch[i].get())"""
        out = np.zeros(max(self.nT - 1, 1), dtype=np.float64)
        n = lib().sa_pt_acceptance(self.p, out.size, _ptr(out))
        check(min(n, 0), "sa_pt_acceptance")
        return out[:n].copy()

    def E_min(self) -> np.ndarray:
        out = np.zeros(64, dtype=np.float32)
        n = lib().sa_pt_E_min(self.p, _ptr(out))
        check(min(n, 0), "sa_pt_E_min")
        return out

    def result(self, spins=True):
        n = lib().sa_pt_steps_done(self.p)
        pairs = np.zeros(max(n, 1), dtype=np.int32)
        us = np.zeros(max(n, 1), dtype=np.float64)
        masks = np.zeros(max(n, 1), dtype=np.uint64)
        self.nT = lib().sa_pt_temperatures(self.p, 0, None)  # update_T / load may change it
        check(min(self.nT, 0), "sa_pt_temperatures")
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


def anneal_geometric(kind, L, T0, alfa, sweeps, nsteps, stride, h=0.0, seed=1234, rng_mode=RNG_COUNTER,
                     precision=0) -> str:
    """annealing<system>::step with geometric cooling (sa_anneal_geometric): the
    driver's printed energy lines."""
    n = lib().sa_anneal_geometric(kind, L, T0, alfa, sweeps, nsteps, stride, h, seed, rng_mode, precision, None, 0)
    check(int(min(n, 0)), "sa_anneal_geometric")
    buf = C.create_string_buffer(int(n) + 1)
    lib().sa_anneal_geometric(kind, L, T0, alfa, sweeps, nsteps, stride, h, seed, rng_mode, precision, buf, len(buf))
    return buf.value.decode()


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

    def add_pt(self, kind, L, T, seed, every, nsteps, h=0.0) -> int:
        """A parallel-tempering instance: len(T) systems, nsteps x (every sweeps + one exchange)."""
        T = np.ascontiguousarray(T, dtype=np.float64)
        i = C.c_int()
        check(lib().sa_batch_add_pt(self.b, kind, L, T.size, T, h, seed, every, nsteps, C.byref(i)), "sa_batch_add_pt")
        self.info[i.value] = (kind, L, T.size)
        return i.value

    def pt_result(self, i: int):
        nT = self.info[i][2]
        n = check(lib().sa_batch_pt_result(self.b, i, 0, None, None, None, None, None), "sa_batch_pt_result")
        pairs = np.zeros(max(n, 1), dtype=np.int32)
        us = np.zeros(max(n, 1), dtype=np.float64)
        masks = np.zeros(max(n, 1), dtype=np.uint64)
        E_min = np.zeros(64, dtype=np.float32)
        acc = np.zeros(nT - 1, dtype=np.float64)
        check(lib().sa_batch_pt_result(self.b, i, n, _ptr(pairs), _ptr(us), _ptr(masks), _ptr(E_min), _ptr(acc)),
              "sa_batch_pt_result")
        return dict(pairs=pairs[:n], us=us[:n], masks=masks[:n], E_min=E_min, acceptance=acc)

    def bytes(self) -> int:
        return lib().sa_batch_bytes(self.b)

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


def make_request(kind, L, mode, nsweeps, seed, nT=1, every=10, T0=None, T1=None, h=0.0, arrival=0.0) -> Request:
    """An anneal (linear T0 -> T1, default 3.0 -> 0.1) or a PT run (geometric ladder, default 0.5 .. 2.0)."""
    if T0 is None:
        T0 = 3.0 if mode == REQ_ANNEAL else 0.5
    if T1 is None:
        T1 = 0.1 if mode == REQ_ANNEAL else 2.0
    return Request(kind, L, mode, nT if mode == REQ_PT else 1, nsweeps, every, T0, T1, h, seed, arrival)


def request_temperatures(r: Request) -> np.ndarray:
    """The temperatures the server runs a request at: per sweep (anneal) or the ladder (PT)."""
    n = check(lib().sa_request_temperatures(C.byref(r), None, 0), "sa_request_temperatures")
    T = np.zeros(n, dtype=np.float64)
    check(lib().sa_request_temperatures(C.byref(r), _ptr(T), n), "sa_request_temperatures")
    return T


def request_bytes(r: Request) -> int:
    """Device bytes the request holds while admitted."""
    return check(lib().sa_request_bytes(C.byref(r)), "sa_request_bytes")


_GOLD = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def digest(words, first_index: int = 0) -> int:
    """XOR over i of splitmix64(w[i] + (first_index + i) * golden) -- sa_digest in numpy."""
    w = np.ascontiguousarray(words, dtype=np.uint64).reshape(-1)
    with np.errstate(over="ignore"):
        idx = (np.arange(w.size, dtype=np.uint64) + np.uint64(first_index)) * _GOLD
        return int(np.bitwise_xor.reduce(_mix64(w + idx))) if w.size else 0


def log_digest(pairs, us, masks) -> int:
    """Chained splitmix64 of (pair, uniform bits, mask) per exchange step."""
    def mix(z):
        z &= 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    h = 0
    ub = np.ascontiguousarray(us, dtype=np.float64).view(np.uint64)
    for p, u, m in zip(pairs, ub, masks):
        h = mix(h ^ (int(p) & 0xFFFFFFFF))
        h = mix(h ^ int(u))
        h = mix(h ^ int(m))
    return h


class Server:
    """Continuous-batching request loop (sa_serve_* of include/sa_hip.h)."""

    def __init__(self, hbm_budget=0, max_active=0, chunk=0, digests=True, prep_threads=0):
        cfg = ServeConfig(hbm_budget, max_active, chunk, 1 if digests else 0, prep_threads)
        s = _vp()
        check(lib().sa_serve_create(C.byref(cfg), C.byref(s)), "sa_serve_create")
        self.s = s

    def clock(self) -> float:
        return lib().sa_serve_clock(self.s)

    def submit(self, r: Request) -> int:
        """Submit; an invalid request is recorded as refused and raises SAError."""
        i = C.c_int64(-1)
        check(lib().sa_serve_submit(self.s, C.byref(r), C.byref(i)), "sa_serve_submit")
        return i.value

    def run(self, n_finished: int = 0):
        check(lib().sa_serve_run(self.s, n_finished), "sa_serve_run")

    def results(self):
        n = lib().sa_serve_results(self.s, 0, 1 << 40, None)
        out = (RequestResult * max(n, 1))()
        lib().sa_serve_results(self.s, 0, n, out)
        return list(out)[:n]

    def stats(self) -> dict:
        st = ServeStats()
        check(lib().sa_serve_get_stats(self.s, C.byref(st)), "sa_serve_get_stats")
        return {k: getattr(st, k) for k, _ in ServeStats._fields_}

    def close(self):
        if getattr(self, "s", None):
            lib().sa_serve_destroy(self.s)
            self.s = None

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
