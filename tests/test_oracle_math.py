"""Pins of the oracle that do not rely on the reference: known-answer tests,
closed forms, invariants and brute force on tiny lattices."""
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import lane_counts

KINDS = [0, 1, 2]


# ---- engines: standard known answers ---------------------------------------------
def test_mt_known_answers(orc):
    # C++ standard [rand.predef]: 10000th output of a default-constructed engine
    g = orc.Mt64(5489)
    v = 0
    for _ in range(10000):
        v = g.next()
    assert v == 9981545732273789042
    g = orc.Mt32(5489)
    for _ in range(10000):
        v = g.next()
    assert v == 4123659995


def test_philox_known_answers(orc):
    # Random123 kat_vectors, philox4x32 10 rounds
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for c, k, want in kat:
        assert list(orc.philox(c, k)) == want


ROCRAND_PROBE = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
struct P : rocrand_device::philox4x32_10_engine {
  static uint4 go(uint4 c, uint2 k) { P p; return p.ten_rounds(c, k); }
};
int main() {
  unsigned a, b, c, d, e, f;
  while (scanf("%x %x %x %x %x %x", &a, &b, &c, &d, &e, &f) == 6) {
    uint4 r = P::go(uint4{a, b, c, d}, uint2{e, f});
    printf("%x %x %x %x\n", r.x, r.y, r.z, r.w);
  }
}
"""


@pytest.mark.skipif(shutil.which("hipcc") is None or not os.path.isdir("/opt/rocm/include/rocrand"),
                    reason="hipcc / rocRAND headers not available")
def test_philox_matches_rocrand_host_implementation(orc):
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        with open(src, "w") as f:
            f.write(ROCRAND_PROBE)
        subprocess.check_call(["hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", src, "-o", exe])
        rng = np.random.default_rng(5)
        rows = rng.integers(0, 2**32, size=(200, 6), dtype=np.uint64)
        inp = "\n".join(" ".join(f"{int(v):x}" for v in r) for r in rows) + "\n"
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
        for r, line in zip(rows, out):
            want = [int(x, 16) for x in line.split()]
            assert list(orc.philox(r[:4], r[4:])) == want


# ---- generate_canonical: closed forms ----------------------------------------------
def test_canonical64_closed_forms(orc):
    assert orc.canonical64(0) == 0.0
    assert orc.canonical64(1 << 63) == 0.5
    assert orc.canonical64(1 << 62) == 0.25
    one_minus = math.nextafter(1.0, 0.0)
    # 2^64-1 rounds up to 2^64 -> 1.0 -> clamped below one
    assert orc.canonical64(2**64 - 1) == one_minus
    # 2^64-1025 rounds to 2^64-2048 (ulp 2048 near 2^64) -> 1 - 2^-53
    assert orc.canonical64(2**64 - 1025) == one_minus
    assert orc.canonical64(2**64 - 3073) == 1.0 - 2.0**-52
    # monotone non-decreasing
    rng = np.random.default_rng(1)
    g = np.sort(rng.integers(0, 2**64, size=20000, dtype=np.uint64))
    u = [orc.canonical64(int(x)) for x in g]
    assert all(a <= b for a, b in zip(u, u[1:]))


def test_canonical32_closed_forms(orc):
    lib = orc.lib()
    assert lib.or_canonical32(0, 0) == 0.0
    assert lib.or_canonical32(0, 1 << 31) == 0.5
    assert lib.or_canonical32(0xFFFFFFFF, 0xFFFFFFFF) == math.nextafter(1.0, 0.0)
    assert lib.or_canonical32(5, 0) == 5 * 2.0**-64


def test_uniform_int_range_and_balance(orc):
    g = orc.Mt32(77)
    for b in (0, 1, 2, 6, 30):
        vals = [g.uniform_int(0, b) for _ in range(3000)]
        assert min(vals) >= 0 and max(vals) <= b
        if b > 0:
            counts = np.bincount(vals, minlength=b + 1)
            assert counts.min() > 3000 / (b + 1) * 0.6


# ---- Boltzmann tables: closed forms --------------------------------------------------
def test_boltz_closed_forms(orc):
    b = orc.boltz(2, 1.0, 0.0)
    assert (b[:4] == 1.0).all() and (b[7:11] == 1.0).all()
    assert b[4] == math.exp(-4.0) and b[5] == math.exp(-8.0) and b[6] == math.exp(-12.0)
    assert (b[:7] == b[7:]).all()
    b = orc.boltz(1, 0.5, 0.0)
    assert (b[:3] == 1.0).all() and b[3] == math.exp(-8.0) and b[4] == math.exp(-16.0)
    assert b[10] == 0.0 and b[15] == 0.0     # unreachable c=0 border entries
    assert b[11] == 1.0 and b[13] == math.exp(-4.0)   # border: j-1 = 2c-5
    b = orc.boltz(2, float("inf"), 0.0)
    assert (b == 1.0).all()


# ---- sweeps: special cases ---------------------------------------------------------------
def _rand_state(kind, L, seed):
    rng = np.random.default_rng(seed)
    N = L * L * (L if kind == 2 else 1)
    r = lambda: rng.integers(0, 2**64, size=N, dtype=np.uint64)  # noqa: E731
    return r(), r(), r(), r()


@pytest.mark.parametrize("kind,L", [(0, 8), (1, 8), (1, 7), (2, 6), (2, 5)])
def test_infinite_temperature_flips_every_spin(orc, kind, L):
    s, jx, jy, jz = _rand_state(kind, L, 3)
    s0 = s.copy()
    words = _rand_state(kind, L, 4)[0]
    orc.sweep(kind, L, s, jx, jy, jz, orc.boltz(kind, float("inf")), words)
    assert (s == ~s0).all()


@pytest.mark.parametrize("kind,L", [(0, 8), (1, 8), (1, 7), (2, 6), (2, 5)])
def test_zero_temperature_ignores_random_numbers(orc, kind, L):
    s, jx, jy, jz = _rand_state(kind, L, 5)
    a, b = s.copy(), s.copy()
    tab = orc.boltz(kind, 1e-3)
    orc.sweep(kind, L, a, jx, jy, jz, tab, _rand_state(kind, L, 6)[0])
    orc.sweep(kind, L, b, jx, jy, jz, tab, _rand_state(kind, L, 7)[0])
    assert (a == b).all()


@pytest.mark.parametrize("kind,L", [(0, 10), (0, 9), (1, 10), (1, 9), (2, 6), (2, 5)])
def test_ferromagnetic_ground_state_energy(orc, kind, L):
    N = L * L * (L if kind == 2 else 1)
    one = np.full(N, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    for s in (np.zeros(N, dtype=np.uint64), one.copy()):
        E, M = orc.measure(kind, L, s, one, one, one)
        nb = {0: 2 * N - L, 1: 2 * N, 2: 3 * N}[kind]
        assert (E == nb).all()
        assert (M == (N if s[0] else -N)).all()


def test_measure_matches_lane_popcounts(orc):
    for kind, L in [(0, 9), (1, 8), (2, 5)]:
        s, jx, jy, jz = _rand_state(kind, L, 11)
        E, M = orc.measure(kind, L, s, jx, jy, jz)
        N = s.size
        assert (M == 2 * lane_counts(s) - N).all()
        idx = np.arange(N)
        x, y, z = idx % L, (idx // L) % L, idx // (L * L)
        west = z * L * L + y * L + (x - 1) % L
        north = z * L * L + ((y - 1) % L) * L + x
        up = ((z - 1) % L) * L * L + y * L + x
        bw = jx ^ s ^ s[west]
        if kind == 0:
            bw = bw[x != 0]
        bn = jy ^ s ^ s[north]
        sat = lane_counts(bw) + lane_counts(bn)
        nb = bw.size + bn.size
        if kind == 2:
            bu = jz ^ s ^ s[up]
            sat = sat + lane_counts(bu)
            nb += bu.size
        assert (E == 2 * sat - nb).all()


def _brute_force_sweep(orc, kind, L, s, jx, jy, jz, T, h, words):
    """Scalar single-spin Metropolis, lane by lane, with the textbook rule
    accept iff u < min(1, exp(-dH/T)), dH = 2 sigma_i (sum_j J_ij sigma_j + h),
    sigma = +1 for a set bit, J_ij = +1 for a set coupling bit.  Same visit order
    and the same uniform per site update as the reference."""
    N = s.size
    order = orc.visit_order(kind, L)
    sig = np.array([[1 if (int(w) >> k) & 1 else -1 for k in range(64)] for w in s], dtype=np.int64)

    def bits(a):
        return np.array([[1 if (int(w) >> k) & 1 else -1 for k in range(64)] for w in a], dtype=np.int64)

    Jx, Jy, Jz = bits(jx), bits(jy), (bits(jz) if kind == 2 else None)
    for k, i in enumerate(order):
        x, y, z = i % L, (i // L) % L, i // (L * L)
        nbrs = []
        if not (kind == 0 and x == 0):
            nbrs.append((z * L * L + y * L + (x - 1) % L, Jx[i]))
        xe = (x + 1) % L
        ie = z * L * L + y * L + xe
        if not (kind == 0 and x == L - 1):
            nbrs.append((ie, Jx[ie]))
        nbrs.append((z * L * L + ((y - 1) % L) * L + x, Jy[i]))
        i_s = z * L * L + ((y + 1) % L) * L + x
        nbrs.append((i_s, Jy[i_s]))
        if kind == 2:
            nbrs.append((((z - 1) % L) * L * L + y * L + x, Jz[i]))
            i_d = ((z + 1) % L) * L * L + y * L + x
            nbrs.append((i_d, Jz[i_d]))
        loc = sum(J * sig[j] for j, J in nbrs)
        dH = 2 * sig[i] * (loc + h)
        u = orc.canonical64(int(words[k]))
        acc = np.array([u < min(1.0, math.exp(-float(d) / T)) for d in dH])
        sig[i] = np.where(acc, -sig[i], sig[i])
    out = np.zeros(N, dtype=np.uint64)
    for i in range(N):
        v = 0
        for kk in range(64):
            if sig[i, kk] > 0:
                v |= 1 << kk
        out[i] = v
    return out


@pytest.mark.parametrize("kind,L,T,h", [(0, 6, 1.1, 0.0), (0, 5, 0.7, 0.3), (1, 6, 1.3, 0.0), (1, 5, 0.9, -0.4),
                                        (2, 4, 1.0, 0.0), (2, 3, 0.8, 0.25)])
def test_sweep_matches_brute_force_metropolis(orc, kind, L, T, h):
    s, jx, jy, jz = _rand_state(kind, L, 21 + L)
    words = _rand_state(kind, L, 31 + L)[0]
    want = _brute_force_sweep(orc, kind, L, s.copy(), jx, jy, jz, T, h, words)
    got = s.copy()
    orc.sweep(kind, L, got, jx, jy, jz, orc.boltz(kind, T, h), words)
    assert (got == want).all()


# ---- swap and the PT mask ---------------------------------------------------------------
def test_swap_is_a_masked_involution(orc):
    rng = np.random.default_rng(9)
    a = rng.integers(0, 2**64, size=100, dtype=np.uint64)
    b = rng.integers(0, 2**64, size=100, dtype=np.uint64)
    m = 0x00FF00F0F0F00001
    a1, b1 = a.copy(), b.copy()
    orc.swap(a1, b1, m)
    mm = np.uint64(m)
    assert ((a1 & mm) == (b & mm)).all() and ((a1 & ~mm) == (a & ~mm)).all()
    assert ((b1 & mm) == (a & mm)).all() and ((b1 & ~mm) == (b & ~mm)).all()
    orc.swap(a1, b1, m)
    assert (a1 == a).all() and (b1 == b).all()


def test_pt_mask_closed_forms(orc):
    Ti, Ti1 = 0.5, 0.8                       # dB = 1/Ti - 1/Ti1 > 0
    lo = np.arange(64, dtype=np.float32) * 4
    hi = lo + 8
    assert orc.pt_mask(lo, hi, Ti, Ti1, 0.999) == 2**64 - 1          # dE > 0: always
    assert orc.pt_mask(lo, lo, Ti, Ti1, 0.999) == 2**64 - 1          # dE = 0: exp(0) = 1 > u
    assert orc.pt_mask(hi, lo, Ti, Ti1, 0.0) == 2**64 - 1            # u = 0 < exp(..) > 0
    assert orc.pt_mask(hi, lo, Ti, Ti1, math.nextafter(1.0, 0.0)) == 0
    # per-lane: dE = -8 -> accept iff u < exp(-8 dB)
    dB = 1 / Ti - 1 / Ti1
    p = math.exp(-8 * dB)
    assert orc.pt_mask(hi, lo, Ti, Ti1, p * 0.999) == 2**64 - 1
    assert orc.pt_mask(hi, lo, Ti, Ti1, p * 1.001) == 0
    Ei = np.zeros(64, dtype=np.float32)
    Ei1 = np.zeros(64, dtype=np.float32)
    Ei[5] = 4.0                               # only lane 5 has dE < 0
    assert orc.pt_mask(Ei, Ei1, Ti, Ti1, 0.999) == (2**64 - 1) ^ (1 << 5)
