"""ctypes bindings of the oracle (oracle/oracle.c) and, when built, of the compiled
reference harness (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
the ``cpu_baseline`` leg of ``bench.py``, always as the checker or as the timed CPU
baseline -- never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

KIND_2P, KIND_2D, KIND_3D = 0, 1, 2

_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile oracle.c into liboracle.so (gcc -O2 -fopenmp, like the reference's makefile)."""
    src = os.path.join(HERE, "oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-std=c99", "-fopenmp", "-fPIC", "-shared", "-o", LIB_PATH, src, "-lm"])
    return LIB_PATH


def build_ref() -> str | None:
    """Build oracle/_ref/libref.so from the reference sources when they are present."""
    subprocess.check_call(["sh", os.path.join(HERE, "build_ref.sh")])
    return REF_PATH if os.path.exists(REF_PATH) else None


class _Mt64(C.Structure):
    _fields_ = [("mt", C.c_uint64 * 312), ("idx", C.c_int)]


class _Mt32(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_mt64_seed.argtypes = [C.POINTER(_Mt64), C.c_uint64]
        L.or_mt64_next.argtypes = [C.POINTER(_Mt64)]
        L.or_mt64_next.restype = C.c_uint64
        L.or_mt64_fill.argtypes = [C.POINTER(_Mt64), _u64p, C.c_int64]
        L.or_mt32_seed.argtypes = [C.POINTER(_Mt32), C.c_uint32]
        L.or_mt32_next.argtypes = [C.POINTER(_Mt32)]
        L.or_mt32_next.restype = C.c_uint32
        L.or_canonical64.argtypes = [C.c_uint64]
        L.or_canonical64.restype = C.c_double
        L.or_canonical32.argtypes = [C.c_uint32, C.c_uint32]
        L.or_canonical32.restype = C.c_double
        L.or_mt32_uniform.argtypes = [C.POINTER(_Mt32)]
        L.or_mt32_uniform.restype = C.c_double
        L.or_mt32_uniform_int.argtypes = [C.POINTER(_Mt32), C.c_int32, C.c_int32]
        L.or_mt32_uniform_int.restype = C.c_int32
        L.or_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.or_counter_word.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint64]
        L.or_counter_word.restype = C.c_uint64
        L.or_boltz_3d.argtypes = [C.c_double, C.c_double, _f64p]
        L.or_boltz_2d.argtypes = [C.c_double, C.c_double, _f64p]
        L.or_sweep.argtypes = [C.c_int, C.c_int, _u64p, _u64p, _u64p, _u64p, _f64p, _u64p]
        L.or_visit_order.argtypes = [C.c_int, C.c_int, _i64p]
        L.or_counter_words.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint64, _u64p]
        L.or_measure.argtypes = [C.c_int, C.c_int, _u64p, _u64p, _u64p, _u64p, _i32p, _i32p]
        L.or_measure_ref_order.argtypes = [C.c_int, C.c_int, _u64p, _u64p, _u64p, _u64p, C.c_double, _f32p]
        L.or_init_J.argtypes = [C.c_int, C.c_int, C.POINTER(_Mt64), _u64p, _u64p, _u64p]
        L.or_swap.argtypes = [C.c_int64, _u64p, _u64p, C.c_uint64]
        L.or_pt_mask.argtypes = [_f32p, _f32p, C.c_double, C.c_double, C.c_double]
        L.or_pt_mask.restype = C.c_uint64
        L.or_pt_run.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_uint64, C.c_int, C.c_int,
                                C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_anneal_run.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int,
                                    C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def nsites(kind: int, L: int) -> int:
    return L * L * (L if kind == KIND_3D else 1)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- engines -------------------------------------------------------------------
class Mt64:
    def __init__(self, seed: int):
        self.st = _Mt64()
        lib().or_mt64_seed(C.byref(self.st), C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF))

    def next(self) -> int:
        return lib().or_mt64_next(C.byref(self.st))

    def fill(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.uint64)
        lib().or_mt64_fill(C.byref(self.st), out, n)
        return out


class Mt32:
    def __init__(self, seed: int):
        self.st = _Mt32()
        lib().or_mt32_seed(C.byref(self.st), seed & 0xFFFFFFFF)

    def next(self) -> int:
        return lib().or_mt32_next(C.byref(self.st))

    def uniform(self) -> float:
        return lib().or_mt32_uniform(C.byref(self.st))

    def uniform_int(self, a: int, b: int) -> int:
        return lib().or_mt32_uniform_int(C.byref(self.st), a, b)


def canonical64(g: int) -> float:
    return lib().or_canonical64(g)


def philox(ctr, key) -> np.ndarray:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(c, k, out)
    return out


def boltz(kind: int, T: float, h: float = 0.0) -> np.ndarray:
    if kind == KIND_3D:
        out = np.zeros(14)
        lib().or_boltz_3d(T, h, out)
    else:
        out = np.zeros(20)
        lib().or_boltz_2d(T, h, out)
    return out


# ---- lattice operations ---------------------------------------------------------
def init_system(kind: int, L: int, seed: int):
    """(jx, jy, jz, s, gen) as main_cpu seeds a system: set_seed, init_J, set_seed,
    init_rand.  The returned engine sits at stream position N, ready for sweeps."""
    N = nsites(kind, L)
    g = Mt64(seed)
    jx = np.zeros(N, dtype=np.uint64)
    jy = np.zeros(N, dtype=np.uint64)
    jz = np.zeros(N, dtype=np.uint64)
    lib().or_init_J(kind, L, C.byref(g.st), jx, jy, jz)
    g = Mt64(seed)
    s = g.fill(N)
    return jx, jy, jz, s, g


def sweep(kind, L, s, jx, jy, jz, table, words):
    lib().or_sweep(kind, L, s, jx, jy, jz, np.ascontiguousarray(table, dtype=np.float64), words)


def visit_order(kind, L) -> np.ndarray:
    out = np.zeros(nsites(kind, L), dtype=np.int64)
    lib().or_visit_order(kind, L, out)
    return out


def counter_words(kind, L, seed, t) -> np.ndarray:
    out = np.zeros(nsites(kind, L), dtype=np.uint64)
    lib().or_counter_words(kind, L, seed, t, out)
    return out


def measure(kind, L, s, jx, jy, jz):
    E = np.zeros(64, dtype=np.int32)
    M = np.zeros(64, dtype=np.int32)
    lib().or_measure(kind, L, s, jx, jy, jz, E, M)
    return E, M


def measure_ref_order(kind, L, s, jx, jy, jz, h=0.0) -> np.ndarray:
    out = np.zeros(64, dtype=np.float32)
    lib().or_measure_ref_order(kind, L, s, jx, jy, jz, h, out)
    return out


def swap(a, b, mask):
    lib().or_swap(a.size, a, b, C.c_uint64(mask))


def pt_mask(Ei, Ei1, Ti, Ti1, u) -> int:
    return lib().or_pt_mask(np.ascontiguousarray(Ei, dtype=np.float32), np.ascontiguousarray(Ei1, dtype=np.float32),
                            Ti, Ti1, u)


def pt_run(kind, L, T, h, seed, sweeps, nsteps, rng_mode=0):
    T = np.ascontiguousarray(T, dtype=np.float64)
    nT = T.size
    N = nsites(kind, L)
    pairs = np.zeros(nsteps, dtype=np.int32)
    us = np.zeros(nsteps, dtype=np.float64)
    masks = np.zeros(nsteps, dtype=np.uint64)
    E = np.zeros(nT * 64, dtype=np.float32)
    spins = np.zeros(nT * N, dtype=np.uint64)
    rc = lib().or_pt_run(kind, L, nT, T, h, seed, sweeps, nsteps, rng_mode, _ptr(pairs), _ptr(us), _ptr(masks),
                         _ptr(E), _ptr(spins))
    assert rc == 0
    return dict(pairs=pairs, us=us, masks=masks, E=E.reshape(nT, 64), spins=spins.reshape(nT, N))


def anneal_run(kind, L, Ts, h, seed, stride, rng_mode=0):
    Ts = np.ascontiguousarray(Ts, dtype=np.float64)
    n = Ts.size
    nrec = (n + stride - 1) // stride
    E = np.zeros(nrec * 64, dtype=np.float32)
    spins = np.zeros(nsites(kind, L), dtype=np.uint64)
    rec = lib().or_anneal_run(kind, L, Ts, n, h, seed, stride, rng_mode, _ptr(E), _ptr(spins))
    return E.reshape(-1, 64)[:rec], spins


# ---- compiled reference (this container only) ---------------------------------------
_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own CPU classes and drivers (oracle/_ref/libref.so)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        R.ref_create.restype = C.c_void_p
        R.ref_destroy.argtypes = [C.c_void_p]
        R.ref_set_seed.argtypes = [C.c_void_p, C.c_long]
        R.ref_init_J.argtypes = [C.c_void_p]
        R.ref_init_rand.argtypes = [C.c_void_p]
        R.ref_init_order.argtypes = [C.c_void_p]
        R.ref_set_T.argtypes = [C.c_void_p, C.c_double]
        R.ref_sweep.argtypes = [C.c_void_p, C.c_int]
        R.ref_measure.argtypes = [C.c_void_p, _f32p]
        R.ref_get_N.argtypes = [C.c_void_p]
        R.ref_get_N.restype = C.c_long
        R.ref_get_state.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]
        R.ref_set_spins.argtypes = [C.c_void_p, C.c_int, _u64p]
        R.ref_peek_word.argtypes = [C.c_void_p, C.c_int]
        R.ref_peek_word.restype = C.c_uint64
        R.ref_swap.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        R.ref_mt64.argtypes = [C.c_uint64, _u64p, C.c_long]
        R.ref_mt32.argtypes = [C.c_uint32, _u32p, C.c_long]
        R.ref_uniform_real64.argtypes = [C.c_uint64, _f64p, C.c_long]
        R.ref_pt_draws.argtypes = [C.c_int, _i32p, _f64p, C.c_long]
        R.ref_save_J.argtypes = [C.c_void_p, C.c_char_p]
        R.ref_load_J.argtypes = [C.c_void_p, C.c_char_p]
        R.ref_pt_run.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_long, C.c_int, C.c_int,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        R.ref_pt_checkpoint.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_long, C.c_int, C.c_int,
                                        C.c_int, C.c_char_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
        R.ref_pt_files.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_long, C.c_int, C.c_int, C.c_int,
                                   C.c_char_p, C.c_char_p, C.c_long]
        R.ref_pt_ground.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_double, C.c_long, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        R.ref_anneal_geometric.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                           C.c_double, C.c_long, C.c_char_p, C.c_long]
        R.ref_time_sweeps.argtypes = [C.c_int, C.c_int, C.c_double, C.c_long, C.c_int]
        R.ref_time_sweeps.restype = C.c_double
        R.ref_time_pt.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, C.c_long, C.c_int, C.c_int, C.c_int]
        R.ref_time_pt.restype = C.c_double
        R.ref_omp_max_threads.restype = C.c_int
        _ref = R
    return _ref


class RefSystem:
    """One reference cpu_2dp / cpu_2d / cpu_3d object."""

    def __init__(self, kind, L, T, h=0.0):
        self.kind, self.L, self.h = kind, L, h
        self.p = ref().ref_create(kind, L, T, h)
        self.N = nsites(kind, L)

    def __del__(self):
        if getattr(self, "p", None):
            ref().ref_destroy(self.p)
            self.p = None

    def seed_like_main(self, seed):
        R = ref()
        R.ref_set_seed(self.p, seed)
        R.ref_init_J(self.p)
        R.ref_set_seed(self.p, seed)
        R.ref_init_rand(self.p)

    def set_T(self, T):
        ref().ref_set_T(self.p, T)

    def sweep(self, n=1):
        ref().ref_sweep(self.p, n)

    def measure(self):
        out = np.zeros(64, dtype=np.float32)
        ref().ref_measure(self.p, out)
        return out

    def state(self):
        N = self.N
        s = np.zeros(N, dtype=np.uint64)
        Jx = np.zeros(2 * N, dtype=np.uint64)
        Jy = np.zeros(2 * N, dtype=np.uint64)
        Jz = np.zeros(2 * N, dtype=np.uint64) if self.kind == KIND_3D else None
        b = np.zeros(14 if self.kind == KIND_3D else 20)
        ref().ref_get_state(self.p, self.kind, _ptr(s), _ptr(Jx), _ptr(Jy), _ptr(Jz), _ptr(b))
        return dict(s=s, Jx=Jx, Jy=Jy, Jz=Jz, boltz=b)

    def set_spins(self, s):
        ref().ref_set_spins(self.p, self.kind, np.ascontiguousarray(s, dtype=np.uint64))

    def peek_word(self) -> int:
        return ref().ref_peek_word(self.p, self.kind)


def ref_pt_run(kind, L, T, h, seed, sweeps, nsteps):
    T = np.ascontiguousarray(T, dtype=np.float64)
    nT = T.size
    N = nsites(kind, L)
    pairs = np.zeros(nsteps, dtype=np.int32)
    us = np.zeros(nsteps, dtype=np.float64)
    masks = np.zeros(nsteps, dtype=np.uint64)
    acc = np.zeros(max(nT - 1, 1), dtype=np.float64)
    E = np.zeros(nT * 64, dtype=np.float32)
    spins = np.zeros(nT * N, dtype=np.uint64)
    ref().ref_pt_run(kind, L, nT, T, h, seed, sweeps, nsteps, _ptr(pairs), _ptr(us), _ptr(masks), _ptr(acc),
                     _ptr(E), _ptr(spins))
    return dict(pairs=pairs, us=us, masks=masks, accept=acc, E=E.reshape(nT, 64), spins=spins.reshape(nT, N))


def ref_pt_checkpoint(kind, L, T, h, seed, sweeps, nsteps1, nsteps2, fname, path):
    """The reference's PT run: nsteps1 steps, parallel_tempering::save to `path`,
    nsteps2 more steps (their pairs, uniforms, masks) and the final spins."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    nT = T.size
    N = nsites(kind, L)
    pairs = np.zeros(max(nsteps2, 1), dtype=np.int32)
    us = np.zeros(max(nsteps2, 1), dtype=np.float64)
    masks = np.zeros(max(nsteps2, 1), dtype=np.uint64)
    spins = np.zeros(nT * N, dtype=np.uint64)
    rc = ref().ref_pt_checkpoint(kind, L, nT, T, h, seed, sweeps, nsteps1, nsteps2, fname.encode(),
                                 str(path).encode(), _ptr(pairs), _ptr(us), _ptr(masks), _ptr(spins))
    assert rc == 0
    return dict(pairs=pairs[:nsteps2], us=us[:nsteps2], masks=masks[:nsteps2], spins=spins.reshape(nT, N))


def ref_pt_files(kind, L, T, h, seed, sweeps, nsteps, stride, pattern) -> str:
    """The reference PT driver's energy files (written under `pattern`) and its
    print_swap text (returned)."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    buf = C.create_string_buffer(1 << 16)
    n = ref().ref_pt_files(kind, L, T.size, T, h, seed, sweeps, nsteps, stride, str(pattern).encode(), buf, len(buf))
    assert n >= 0
    return buf.value.decode()


def ref_pt_ground(kind, L, T, h, seed, sweeps, nsteps, rounds, cap=16):
    """The reference's ground-state loop pieces on one PT driver (rounds of steps,
    get_E_min, update_T, reset; then steps once more)."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    N = L ** (3 if kind == 2 else 2)
    nT_out = np.zeros(rounds, dtype=np.int32)
    acc = np.zeros(rounds * cap, dtype=np.float64)
    E_min = np.zeros(rounds * 64, dtype=np.float32)
    T_out = np.zeros(rounds * cap, dtype=np.float64)
    spins = np.zeros(cap * N, dtype=np.uint64)
    E = np.zeros(cap * 64, dtype=np.float32)
    n = ref().ref_pt_ground(kind, L, T.size, T, h, seed, sweeps, nsteps, rounds, cap, nT_out.ctypes.data,
                            acc.ctypes.data, E_min.ctypes.data, T_out.ctypes.data, spins.ctypes.data, E.ctypes.data)
    assert n > 0, "over capacity"
    return dict(nT=nT_out, acc=acc.reshape(rounds, cap), E_min=E_min.reshape(rounds, 64),
                T_rounds=[T_out[r * cap:r * cap + int(nT_out[r])] for r in range(rounds)],
                spins=spins.reshape(cap, N)[:n], E=E.reshape(cap, 64)[:n])


def ref_anneal_geometric(kind, L, T0, alfa, sweeps, nsteps, stride, h, seed) -> str:
    buf = C.create_string_buffer(1 << 22)
    n = ref().ref_anneal_geometric(kind, L, T0, alfa, sweeps, nsteps, stride, h, seed, buf, len(buf))
    assert n >= 0
    return buf.value.decode()
