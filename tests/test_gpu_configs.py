"""The benchmark configurations at their stated sizes, on the HIP path.

Replay mode (the reference's own mt19937_64 stream) against fixtures written by
the compiled reference (tests/golden/config_vectors.json, make_golden.py --configs):
  * 2D periodic L=64, linear anneal 3.0 -> 0.1 over 1,000 sweeps, energies every
    10 sweeps;
  * 3D L=64 at T=1.0, 10,000 sweeps: energies and spins every 1,000 sweeps;
  * 3D L=128 parallel tempering, 32 temperatures geometric 0.5 .. 2.0, 10 sweeps per
    exchange: pairs, uniforms, masks, all 32 x 64 energies and every spin array.
Counter mode at the benchmark's own kernel configuration (fused sweep, two systems
per workgroup, bands of eight rows) against the oracle fed the same Philox words.
Bar: bit-exact."""
import json
import os

import numpy as np
import pytest

from helpers import f32, f64, sha, words

pytestmark = pytest.mark.gpu

CONFIGS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config_vectors.json")


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


@pytest.fixture(scope="module")
def cfg():
    with open(CONFIGS) as f:
        return json.load(f)


def test_config_anneal2d_L64_1000_sweeps_matches_reference(sa, cfg):
    c = cfg["anneal2d_L64"]
    say("2d periodic L=64 linear anneal, 1000 sweeps, replay")
    E, s = sa.anneal_run(c["kind"], c["L"], np.array(c["T"]), 0.0, c["seed"], c["stride"], sa.RNG_REPLAY)
    assert len(E) == len(c["E"]) == 100
    for k, (a, b) in enumerate(zip(E, c["E"])):
        assert (a == f32(b)).all(), k
    assert sha(s) == c["s_sha"]


def test_config_fixed3d_L64_10000_sweeps_matches_reference(sa, cfg):
    c = cfg["fixed3d_L64"]
    say("3d L=64 T=1, 10000 sweeps, replay")
    lad = sa.Ladder(c["kind"], c["L"], 1)
    lad.set_rng(sa.RNG_REPLAY)
    lad.seed_like_main(c["seed"])
    lad.set_temperature(0, c["T"], 0.0)
    for k in range(c["sweeps"] // c["every"]):
        lad.sweep(c["every"])
        E, _ = lad.measure()
        assert (E[0].astype(np.float32) == f32(c["E"][k])).all(), k
        assert sha(lad.spins(0)) == c["s_sha"][k], k
        say("  ", (k + 1) * c["every"], "sweeps ok")
    lad.close()


def tuning(sa, handle):
    import ctypes as C

    out = {}
    for name, key in (("kernel", sa.TUNE_SWEEP3D_KERNEL), ("mb", sa.TUNE_SYS_PER_THREAD),
                      ("th", sa.TUNE_ROWS_PER_TILE)):
        v = C.c_int()
        sa.check(sa.lib().sa_ladder_get_tuning(handle, key, C.byref(v)), "get_tuning")
        out[name] = v.value
    return out


def test_config_pt3d_L128_32_temperatures_matches_reference(sa, cfg):
    c = cfg["pt3d_L128"]
    say("3d L=128 PT, 32 temperatures, 10 sweeps per exchange, replay")
    pt = sa.PT(c["kind"], c["L"], np.array(c["T"]), c["h"], c["seed"], c["sweeps"], sa.RNG_REPLAY)
    # the kernel the benchmark times: fused sweep, two systems per workgroup, 8-row bands
    assert tuning(sa, pt.ladder_handle()) == dict(kernel=2, mb=2, th=8)
    pt.step(c["nsteps"])
    r = pt.result()
    assert list(r["pairs"]) == c["pairs"]
    assert (r["us"] == f64(c["us"])).all()
    assert (r["masks"] == words(c["masks"])).all()
    for j, e in enumerate(c["E"]):
        assert (r["E"][j] == f32(e)).all(), j
    assert [sha(s) for s in r["spins"]] == c["spins_sha"]
    pt.close()


def test_bench_kernel_counter_pt_L128_32_temperatures_matches_oracle(sa, orc):
    """The benchmark's workload and kernel instantiation (counter RNG, fused sweep,
    MB=2, TH=8): one exchange step of 2 sweeps against the oracle."""
    kind, L, nT, sweeps, seed = 2, 128, 32, 2, 1234
    T = np.array([0.5 * 4.0 ** (k / (nT - 1)) for k in range(nT)])
    say("bench kernel: counter PT L=128 x 32, 1 step of 2 sweeps")
    pt = sa.PT(kind, L, T, 0.0, seed, sweeps, sa.RNG_COUNTER)
    assert tuning(sa, pt.ladder_handle()) == dict(kernel=2, mb=2, th=8)
    pt.step(1)
    a = pt.result()
    pt.close()
    b = orc.pt_run(kind, L, T, 0.0, seed, sweeps, 1, rng_mode=1)
    for k in ("pairs", "us", "masks", "E"):
        assert (a[k] == b[k]).all(), k
    for j in range(nT):
        assert (a["spins"][j] == b["spins"][j]).all(), j


def test_counter_two_systems_forced_mb2_th8_L128_matches_oracle(sa, orc):
    """Two systems at different temperatures forced onto the benchmark's kernel
    configuration (two systems per workgroup, 8-row bands), 2 sweeps + measure."""
    kind, L, seed = 2, 128, 99
    say("counter L=128, 2 systems forced MB=2 TH=8")
    lad = sa.Ladder(kind, L, 2)
    lad.set_rng(sa.RNG_COUNTER)
    lad.seed_like_main(seed)
    Ts = [0.7, 1.6]
    for j, t in enumerate(Ts):
        lad.set_temperature(j, t)
    lad.tune(sa.TUNE_SYS_PER_THREAD, 2)
    lad.tune(sa.TUNE_ROWS_PER_TILE, 8)
    assert tuning(sa, lad.h) == dict(kernel=2, mb=2, th=8)
    jx, jy, jz, s0, _ = orc.init_system(kind, L, seed)
    ref = [s0.copy(), s0.copy()]
    lad.sweep(2)
    for t in range(2):
        w = orc.counter_words(kind, L, seed, t)
        for j in range(2):
            orc.sweep(kind, L, ref[j], jx, jy, jz, orc.boltz(kind, Ts[j]), w)
    E, M = lad.measure()
    for j in range(2):
        assert (lad.spins(j) == ref[j]).all(), j
        e, m = orc.measure(kind, L, ref[j], jx, jy, jz)
        assert (E[j] == e).all() and (M[j] == m).all(), j
    lad.close()
