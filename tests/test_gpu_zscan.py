"""The z-parallel 3D half-sweep (k_sweep3d_zscan: each column's z walk split into
segments whose per-plane bit maps are composed by a scan across lanes) against the
oracle's in-order walk, bit for bit: counter and replay streams, symmetric and
general tables (h != 0), every segment length the kernel has (L/G = 2..8)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


def zscan_ladder(sa, L, nsys, seed, rng, Ts, hs):
    lad = sa.Ladder(2, L, nsys)
    lad.set_rng(rng)
    lad.seed_like_main(seed)
    for j in range(nsys):
        lad.set_temperature(j, Ts[j], hs[j])
    lad.tune(sa.TUNE_SWEEP3D_KERNEL, 3)
    assert lad.tuning(sa.TUNE_SWEEP3D_KERNEL) == 3
    return lad


# L: G x S = 4: 2x2, 6: 2x3, 8: 4x2, 10: 2x5, 12: 4x3, 14: 2x7, 16: 8x2, 24: 8x3, 32: 16x2,
# 40: 8x5, 48: 16x3, 56: 8x7, 64: 32x2, 96: 32x3, 128: 32x4, 256: 64x4
@pytest.mark.parametrize("L", [4, 6, 8, 10, 12, 14, 16, 24, 32, 40, 48, 56, 64, 96, 128, 256])
def test_zscan_counter_matches_oracle(sa, orc, L):
    nsys, nsw = (2, 3) if L <= 128 else (1, 1)
    print("[case] zscan counter L", L, "nsys", nsys, flush=True)
    seed = 4242 + L
    Ts = [0.55, 1.4][:nsys]
    hs = [0.0, 0.25][:nsys]
    lad = zscan_ladder(sa, L, nsys, seed, sa.RNG_COUNTER, Ts, hs)
    jx, jy, jz, s0, _ = orc.init_system(2, L, seed)
    ref = [s0.copy() for _ in range(nsys)]
    lad.sweep(nsw)
    for t in range(nsw):
        w = orc.counter_words(2, L, seed, t)
        for j in range(nsys):
            orc.sweep(2, L, ref[j], jx, jy, jz, orc.boltz(2, Ts[j], hs[j]), w)
    for j in range(nsys):
        assert (lad.spins(j) == ref[j]).all(), j
    lad.close()


@pytest.mark.parametrize("L", [6, 16, 64])
def test_zscan_replay_matches_oracle(sa, orc, L):
    print("[case] zscan replay L", L, flush=True)
    seed, nsw = 99 + L, 3
    lad = zscan_ladder(sa, L, 1, seed, sa.RNG_REPLAY, [0.8], [0.0])
    jx, jy, jz, s, g = orc.init_system(2, L, seed)
    lad.sweep(nsw)
    for _ in range(nsw):
        orc.sweep(2, L, s, jx, jy, jz, orc.boltz(2, 0.8), g.fill(s.size))
    assert (lad.spins(0) == s).all()
    lad.close()
