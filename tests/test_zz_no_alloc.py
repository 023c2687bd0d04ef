"""sa_measure_noalloc (the measurement with no allocation, counters in a fixed
device buffer) against sa_ladder_measure on the same states, through the C ABI:
two calls in a row, the second on a smaller input, each checked bit for bit.
The test's device arrays and stream come from the HIP runtime the library itself
links (/opt/rocm's libamdhip64, through ctypes), not from torch: torch ships a HIP
runtime of its own, and a second runtime in the process could not initialise the
device after the other GPU tests had run."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def ladder_state(sa, kind, L, nsys, seed, sweeps):
    lad = sa.Ladder(kind, L, nsys)
    lad.set_rng(sa.RNG_COUNTER)
    lad.seed_like_main(seed)
    for j in range(nsys):
        lad.set_temperature(j, 0.6 + 0.3 * j, 0.1 * (j % 2) if kind == sa.KIND_2P else 0.0)
    lad.sweep(sweeps)
    E, M = lad.measure()
    jx, jy, jz = lad.couplings()
    s = np.stack([lad.spins(j) for j in range(nsys)])
    lad.close()
    return E, M, s, jx, jy, jz


class Hip:
    """hipMalloc / hipMemcpy / streams of the runtime libsa_hip.so is linked against."""
    H2D, D2H = 1, 2

    def __init__(self):
        self.rt = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7"))
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.rt.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.live = []

    def ok(self, rc, what):
        assert rc == 0, f"{what}: hipError {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)), "hipMalloc")
        self.live.append(p.value)
        self.ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D), "hipMemcpy H2D")
        return p.value

    def download(self, p, like):
        out = np.empty_like(like)
        self.ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def noalloc(sa, hip, stream, kind, L, nsys, s, jx, jy, jz):
    ds, djx, djy = hip.upload(s.reshape(-1)), hip.upload(jx), hip.upload(jy)
    djz = hip.upload(jz) if jz is not None else None
    # outputs prefilled with garbage: every entry must be written
    junk = np.full(nsys * 64, -7, dtype=np.int32)
    dE, dM = hip.upload(junk), hip.upload(junk)
    rc = sa.lib().sa_measure_noalloc(kind, L, nsys, ds, djx, djy, djz, dE, dM, stream)
    sa.check(rc, "sa_measure_noalloc")
    hip.ok(hip.rt.hipStreamSynchronize(stream), "hipStreamSynchronize")
    E, M = hip.download(dE, junk), hip.download(dM, junk)
    hip.free()
    return E.reshape(nsys, 64), M.reshape(nsys, 64)


@pytest.mark.parametrize("cases", [
    [(2, 16, 6, 11), (2, 8, 2, 12)],       # 3D: 6 systems (4 per block row), then 2 (one per row)
    [(1, 64, 9, 21), (0, 10, 1, 22)],      # 2D periodic, then open-x 2p with one system
    [(2, 35, 3, 31), (1, 9, 2, 32)],       # odd L
])
def test_noalloc_measure_equals_ladder_measure_twice(cases):
    import simulated_annealing_amd as sa

    sa.lib()
    hip = Hip()
    st = C.c_void_p()
    hip.ok(hip.rt.hipStreamCreate(C.byref(st)), "hipStreamCreate")
    try:
        # first call on a stream of its own, the second (smaller) on the null stream
        for (kind, L, nsys, seed), stream in zip(cases, (st.value, None)):
            say("noalloc measure kind", kind, "L", L, "nsys", nsys, "stream", "own" if stream else "null")
            E0, M0, s, jx, jy, jz = ladder_state(sa, kind, L, nsys, seed, 3)
            E1, M1 = noalloc(sa, hip, stream, kind, L, nsys, s, jx, jy, jz if kind == sa.KIND_3D else None)
            assert (E1 == E0).all() and (M1 == M0).all(), (kind, L, nsys)
    finally:
        hip.free()
        hip.rt.hipStreamDestroy(st)


def test_noalloc_refuses_more_systems_than_its_buffer():
    import simulated_annealing_amd as sa

    L = sa.lib()
    rc = L.sa_measure_noalloc(sa.KIND_2D, 8, 65, 1, 1, 1, None, 1, 1, None)
    assert rc == -1 and b"65" in L.sa_last_error()
