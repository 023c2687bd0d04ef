"""Parity of the HIP path (through the C ABI) with the reference.

Replay mode (the reference's own mt19937_64 stream) is checked against the
golden fixtures written from the compiled reference; counter mode (Philox words)
against the oracle fed the same words.  Bar: bit-exact spin words, integer
energies and swap decisions."""
import numpy as np
import pytest

from helpers import f32, f64, sha, words

pytestmark = pytest.mark.gpu

KIND_NAME = {0: "2p", 1: "2d", 2: "3d"}


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


def device_cases(golden):
    return [c for c in golden["sweeps"] if c["L"] % 2 == 0 and c["L"] >= 4]


# ---- replay mode vs the compiled reference's fixtures ---------------------------------
def test_replay_sweeps_match_reference_fixtures(sa, golden):
    for c in device_cases(golden):
        kind, L = c["kind"], c["L"]
        say("replay sweeps", KIND_NAME[kind], "L", L, "T", c["T"], "h", c["h"])
        lad = sa.Ladder(kind, L, 1)
        lad.set_rng(sa.RNG_REPLAY)
        lad.seed_like_main(c["seed"])
        lad.set_temperature(0, c["T"], c["h"])
        assert sha(lad.spins(0)) == c["s0_sha"]
        jx, jy, jz = lad.couplings()
        assert sha(jx) == c["jx_sha"] and sha(jy) == c["jy_sha"]
        if kind == 2:
            assert sha(jz) == c["jz_sha"]
        E, M = lad.measure()
        got = (E[0].astype(np.float64) + c["h"] * M[0].astype(np.float64)).astype(np.float32)
        assert (got == f32(c["E0"])).all()
        for k in range(c["sweeps"]):
            lad.sweep(1)
            E, M = lad.measure()
            got = (E[0].astype(np.float64) + c["h"] * M[0].astype(np.float64)).astype(np.float32)
            assert (got == f32(c["E"][k])).all(), (c["kind"], L, k)
        s = lad.spins(0)
        assert sha(s) == c["s_sha"]
        if "s" in c:
            assert (s == words(c["s"])).all()
        lad.close()


def test_replay_multi_sweep_call_matches_single_calls(sa, golden):
    c = [c for c in device_cases(golden) if c["kind"] == 2][0]
    say("replay sweep(n) vs n x sweep(1)", c["L"])
    lad = sa.Ladder(2, c["L"], 1)
    lad.seed_like_main(c["seed"])
    lad.set_temperature(0, c["T"], c["h"])
    lad.sweep(c["sweeps"])
    assert sha(lad.spins(0)) == c["s_sha"]


def test_boltz_tables_match_reference_bits(sa, golden):
    for t in golden["tables"]:
        say("boltz", t)
        lad = sa.Ladder(t["kind"], 4, 1)
        b = lad.set_temperature(0, t["T"], t["h"])
        assert (b.view(np.uint64) == f64(t["boltz"]).view(np.uint64)).all()
        lad.close()


def test_schedule_anneal_matches_reference_fixture(sa, golden):
    c = golden["schedule"]
    say("anneal schedule", c["kind"], c["L"])
    E, s = sa.anneal_run(c["kind"], c["L"], np.array(c["T"]), 0.0, c["seed"], c["stride"], sa.RNG_REPLAY)
    assert len(E) == len(c["E"])
    for a, b in zip(E, c["E"]):
        assert (a == f32(b)).all()
    assert sha(s) == c["s_sha"]


def test_pt_replay_matches_reference_fixtures(sa, golden):
    for c in golden["pt"]:
        if c["L"] % 2:
            continue
        say("pt replay", KIND_NAME[c["kind"]], "L", c["L"], "nT", len(c["T"]), "h", c["h"])
        r = sa.pt_run(c["kind"], c["L"], np.array(c["T"]), c["h"], c["seed"], c["sweeps"], c["nsteps"],
                      sa.RNG_REPLAY)
        assert list(r["pairs"]) == c["pairs"]
        assert (r["us"] == f64(c["us"])).all()
        assert (r["masks"] == words(c["masks"])).all()
        for j, e in enumerate(c["E"]):
            assert (r["E"][j] == f32(e)).all()
        assert [sha(s) for s in r["spins"]] == c["spins_sha"]


# ---- counter mode vs the oracle fed the same Philox words -------------------------------
@pytest.mark.parametrize("kind,L,nsys,nsw", [(2, 8, 1, 6), (2, 16, 3, 4), (2, 32, 4, 3), (1, 16, 2, 8),
                                             (1, 64, 4, 5), (0, 16, 2, 8), (0, 30, 1, 6), (2, 6, 2, 5)])
def test_counter_sweeps_match_oracle(sa, orc, kind, L, nsys, nsw):
    say("counter sweeps", KIND_NAME[kind], "L", L, "nsys", nsys, "sweeps", nsw)
    seed = 777 + L
    lad = sa.Ladder(kind, L, nsys)
    lad.set_rng(sa.RNG_COUNTER)
    lad.seed_like_main(seed)
    Ts = [0.6 + 0.7 * j for j in range(nsys)]
    hs = [0.0 if j % 2 == 0 else 0.3 for j in range(nsys)]
    for j in range(nsys):
        lad.set_temperature(j, Ts[j], hs[j])
    jx, jy, jz, s0, _ = orc.init_system(kind, L, seed)
    ref = [s0.copy() for _ in range(nsys)]
    for t in range(nsw):
        lad.sweep(1)
        w = orc.counter_words(kind, L, seed, t)
        for j in range(nsys):
            orc.sweep(kind, L, ref[j], jx, jy, jz, orc.boltz(kind, Ts[j], hs[j]), w)
        E, M = lad.measure()
        for j in range(nsys):
            e, m = orc.measure(kind, L, ref[j], jx, jy, jz)
            assert (E[j] == e).all() and (M[j] == m).all(), (t, j)
    for j in range(nsys):
        assert (lad.spins(j) == ref[j]).all(), j
    assert lad.sweep_index == nsw


def test_counter_3d_L128_two_sweeps_match_oracle(sa, orc):
    kind, L = 2, 128
    say("counter 3d L128 two sweeps")
    lad = sa.Ladder(kind, L, 1)
    lad.set_rng(sa.RNG_COUNTER)
    lad.seed_like_main(1234)
    lad.set_temperature(0, 0.5, 0.0)
    jx, jy, jz, s, _ = orc.init_system(kind, L, 1234)
    lad.sweep(2)
    for t in range(2):
        orc.sweep(kind, L, s, jx, jy, jz, orc.boltz(kind, 0.5), orc.counter_words(kind, L, 1234, t))
    assert (lad.spins(0) == s).all()


def test_pt_counter_matches_oracle(sa, orc):
    for kind, L, nT, sweeps, nsteps in [(2, 8, 8, 4, 12), (1, 16, 5, 3, 20), (0, 12, 4, 2, 15)]:
        say("pt counter", KIND_NAME[kind], L, nT)
        T = np.array([0.5 * 4 ** (k / (nT - 1)) for k in range(nT)])
        a = sa.pt_run(kind, L, T, 0.0, 4321, sweeps, nsteps, sa.RNG_COUNTER)
        b = orc.pt_run(kind, L, T, 0.0, 4321, sweeps, nsteps, rng_mode=1)
        for k in ("pairs", "us", "masks", "E", "spins"):
            assert (a[k] == b[k]).all(), k


# ---- measure / swap on arbitrary states ------------------------------------------------
@pytest.mark.parametrize("kind,L", [(0, 20), (1, 20), (2, 10), (2, 64)])
def test_measure_and_swap_on_random_states(sa, orc, kind, L):
    say("measure/swap", KIND_NAME[kind], L)
    rng = np.random.default_rng(L)
    N = L * L * (L if kind == 2 else 1)
    r = lambda: rng.integers(0, 2**64, size=N, dtype=np.uint64)  # noqa: E731
    jx, jy, jz, a, b = r(), r(), r(), r(), r()
    lad = sa.Ladder(kind, L, 3)
    lad.set_couplings(jx, jy, jz if kind == 2 else None)
    lad.set_spins(0, a)
    lad.set_spins(2, b)
    E, M = lad.measure()
    for j, s in ((0, a), (2, b)):
        e, m = orc.measure(kind, L, s, jx, jy, jz)
        assert (E[j] == e).all() and (M[j] == m).all()
    mask = 0xF0F0_0000_FFFF_0001
    lad.swap(0, 2, mask)
    orc.swap(a, b, mask)
    assert (lad.spins(0) == a).all() and (lad.spins(2) == b).all()


# ---- size-independent properties at the benchmark size ----------------------------------
def test_infinite_temperature_flips_every_spin_L128(sa):
    say("T=inf flips, 3d L128 x 2 systems")
    lad = sa.Ladder(2, 128, 2)
    lad.set_rng(sa.RNG_COUNTER)
    lad.seed_like_main(5)
    lad.set_temperature(0, float("inf"))
    lad.set_temperature(1, float("inf"))
    s0 = lad.spins(1)
    lad.sweep(1)
    assert (lad.spins(1) == ~s0).all()
    lad.sweep(1)
    assert (lad.spins(1) == s0).all()


def test_ladder_slot_equals_lone_system_L128(sa):
    """A system's result does not depend on the other slots of its ladder."""
    say("ladder slot vs alone, 3d L128")
    T = [0.5, 0.9, 1.4, 2.0]
    big = sa.Ladder(2, 128, 4)
    big.set_rng(sa.RNG_COUNTER)
    big.seed_like_main(99)
    for j, t in enumerate(T):
        big.set_temperature(j, t)
    big.sweep(3)
    lone = sa.Ladder(2, 128, 1)
    lone.set_rng(sa.RNG_COUNTER)
    lone.seed_like_main(99)
    lone.set_temperature(0, T[2])
    lone.sweep(3)
    assert (big.spins(2) == lone.spins(0)).all()
    E1, M1 = big.measure()
    E2, M2 = lone.measure()
    assert (E1[2] == E2[0]).all() and (M1[2] == M2[0]).all()


def test_ferromagnetic_ground_state_L128(sa):
    say("ground state energy 3d L128")
    L = 128
    N = L**3
    one = np.full(N, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    lad = sa.Ladder(2, L, 1)
    lad.set_couplings(one, one, one)
    lad.set_spins(0, np.zeros(N, dtype=np.uint64))
    E, M = lad.measure()
    assert (E == 3 * N).all() and (M == -N).all()
    # at T -> 0 a ground state stays put
    lad.set_temperature(0, 1e-3)
    lad.set_rng(sa.RNG_COUNTER)
    lad.sweep(2)
    assert (lad.spins(0) == 0).all()


def test_unsupported_geometry_is_an_error(sa):
    with pytest.raises(sa.SAError):
        sa.Ladder(2, 2, 1)
    with pytest.raises(sa.SAError):
        sa.Ladder(1, 2, 1)
    b = sa.Batch()
    with pytest.raises(sa.SAError):
        b.add(2, 2, 1, 3, T=np.ones(1))  # L < 3 in a batch
    b.close()


# ---- every 3D kernel variant against the oracle ----------------------------------------
@pytest.mark.parametrize("L,nsys,kernel,th,mb,xcd", [
    (8, 1, 2, 8, 1, -1), (8, 2, 2, 4, 2, 0), (16, 4, 2, 16, 2, -1), (16, 2, 2, 8, 1, 1), (32, 4, 2, 8, 2, 1),
    (32, 2, 1, 0, 2, -1), (64, 2, 2, 16, 1, 0), (16, 4, 2, 4, 1, 1), (4, 2, 2, 4, 1, -1), (32, 6, 2, 4, 2, -1)])
@pytest.mark.parametrize("mode", [0, 1])
def test_3d_kernel_variants_match_oracle(sa, orc, L, nsys, kernel, th, mb, xcd, mode):
    say("3d variant", "L", L, "nsys", nsys, "kernel", kernel, "th", th, "mb", mb, "xcd", xcd, "mode", mode)
    kind, seed, nsw = 2, 31 + L, 3
    lad = sa.Ladder(kind, L, nsys)
    lad.set_rng(mode)
    lad.tune(sa.TUNE_SWEEP3D_KERNEL, kernel)
    lad.tune(sa.TUNE_SYS_PER_THREAD, mb)
    lad.tune(sa.TUNE_ROWS_PER_TILE, th)
    lad.tune(sa.TUNE_XCD_MAP, xcd)
    lad.seed_like_main(seed)
    Ts = [0.4 + 0.5 * j for j in range(nsys)]
    hs = [0.0 if j % 3 else 0.25 for j in range(nsys)] if mode == 1 else [0.0] * nsys
    for j in range(nsys):
        lad.set_temperature(j, Ts[j], hs[j])
    assert lad.tuning(sa.TUNE_SWEEP3D_KERNEL) == kernel
    if kernel == 2:
        assert lad.tuning(sa.TUNE_ROWS_PER_TILE) == th
    jx, jy, jz, s0, g = orc.init_system(kind, L, seed)
    ref = [s0.copy() for _ in range(nsys)]
    for t in range(nsw):
        w = orc.counter_words(kind, L, seed, t) if mode == 1 else g.fill(L**3)
        for j in range(nsys):
            orc.sweep(kind, L, ref[j], jx, jy, jz, orc.boltz(kind, Ts[j], hs[j]), w)
    lad.sweep(nsw)
    for j in range(nsys):
        assert (lad.spins(j) == ref[j]).all(), j
    E, M = lad.measure()
    for j in range(nsys):
        e, m = orc.measure(kind, L, ref[j], jx, jy, jz)
        assert (E[j] == e).all() and (M[j] == m).all()


def test_fused_and_half_sweep_agree_at_L128(sa):
    say("fused vs half-sweep, 3d L128 x 4 systems")
    out = []
    for kernel in (1, 2):
        lad = sa.Ladder(2, 128, 4)
        lad.set_rng(sa.RNG_COUNTER)
        lad.tune(sa.TUNE_SWEEP3D_KERNEL, kernel)
        lad.seed_like_main(17)
        for j, t in enumerate([0.5, 0.8, 1.2, 2.0]):
            lad.set_temperature(j, t)
        lad.sweep(4)
        out.append([lad.spins(j) for j in range(4)])
        lad.close()
    for a, b in zip(*out):
        assert (a == b).all()


def test_measure_systems_subset_matches_full_measure(sa):
    say("measure subset, 3d L32 x 6 systems")
    lad = sa.Ladder(2, 32, 6)
    lad.set_rng(sa.RNG_COUNTER)
    lad.seed_like_main(8)
    for j in range(6):
        lad.set_temperature(j, 0.5 + 0.3 * j)
    lad.sweep(3)
    E, M = lad.measure()
    for subset in ([0], [5, 1], [2, 3, 4], [0, 1, 2, 3, 4, 5], [4, 4]):
        Es, Ms = lad.measure_systems(subset)
        for q, j in enumerate(subset):
            assert (Es[q] == E[j]).all() and (Ms[q] == M[j]).all()
    with pytest.raises(sa.SAError):
        lad.measure_systems([6])


def test_ladders_swept_from_concurrent_threads(sa):
    """Calls on different ladders may come from different threads (the reference's
    parallel_tempering sweeps systems from OpenMP threads): four ladders swept
    concurrently end exactly as when swept one after another."""
    import threading

    cases = [(2, 16, 2, 11), (1, 32, 3, 12), (0, 24, 1, 13), (2, 8, 4, 14)]

    def make(kind, L, nsys, seed):
        lad = sa.Ladder(kind, L, nsys)
        lad.set_rng(sa.RNG_COUNTER)
        lad.seed_like_main(seed)
        for j in range(nsys):
            lad.set_temperature(j, 0.6 + 0.4 * j)
        return lad

    def work(lad, out, q):
        for _ in range(20):
            lad.sweep(1)
            lad.measure()
        out[q] = [lad.spins(j) for j in range(lad.nsys)]

    seq = {}
    for q, c in enumerate(cases):
        work(make(*c), seq, q)
    par = {}
    lads = [make(*c) for c in cases]
    ths = [threading.Thread(target=work, args=(lad, par, q)) for q, lad in enumerate(lads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for q in range(len(cases)):
        for a, b in zip(seq[q], par[q]):
            assert (a == b).all(), cases[q]


def test_geometric_anneal_driver_text_matches_reference(sa, golden):
    """annealing<system>::step with geometric cooling on the device (sa_anneal_geometric),
    replay stream, against the compiled reference driver's printed lines
    (src/annealing.hpp:67-100): byte for byte at the fixture's stream precision (9)."""
    for c in golden["anneal"]:
        say("geometric anneal", c["kind"], c["L"], c["T0"], c["alfa"])
        txt = sa.anneal_geometric(c["kind"], c["L"], c["T0"], c["alfa"], c["sweeps"], c["nsteps"], c["stride"],
                                  0.0, c["seed"], sa.RNG_REPLAY, precision=9)
        assert txt == c["text"]


@pytest.mark.parametrize("mb", [1, 2])
def test_replay_fused_L128_matches_oracle(sa, orc, mb):
    """The fused sweep at L=128 (rows are whole waves: the parity-specialised walk)
    fed the reference's own mt19937_64 stream, against the oracle: two sweeps of two
    systems, one or two systems per thread."""
    kind, L, nsys, seed = 2, 128, 2, 777
    say("replay fused L=128", "mb", mb)
    lad = sa.Ladder(kind, L, nsys)
    lad.set_rng(sa.RNG_REPLAY)
    lad.tune(sa.TUNE_SWEEP3D_KERNEL, 2)
    lad.tune(sa.TUNE_SYS_PER_THREAD, mb)
    lad.tune(sa.TUNE_ROWS_PER_TILE, 8)
    lad.seed_like_main(seed)
    Ts = [0.7, 1.6]
    for j in range(nsys):
        lad.set_temperature(j, Ts[j])
    jx, jy, jz, s0, g = orc.init_system(kind, L, seed)
    ref = [s0.copy() for _ in range(nsys)]
    for _ in range(2):
        w = g.fill(L ** 3)
        for j in range(nsys):
            orc.sweep(kind, L, ref[j], jx, jy, jz, orc.boltz(kind, Ts[j], 0.0), w)
    lad.sweep(2)
    for j in range(nsys):
        assert (lad.spins(j) == ref[j]).all(), j
    lad.close()


# ---- odd L: the wrap neighbours share a colour --------------------------------------------
@pytest.mark.parametrize("kind,L,nsys", [(2, 3, 1), (2, 5, 2), (2, 7, 1), (2, 9, 3), (2, 15, 1), (1, 3, 1),
                                         (1, 5, 1), (1, 9, 2), (1, 21, 1), (0, 5, 1), (0, 7, 2), (0, 13, 1),
                                         # 3D: the lane pairs span two waves (L >= 35), the largest
                                         # served 3D size; 2D: the largest served size
                                         (2, 35, 1), (2, 63, 2), (1, 129, 1), (0, 255, 1)])
@pytest.mark.parametrize("mode", [0, 1])
def test_odd_L_matches_oracle(sa, orc, kind, L, nsys, mode):
    """Odd L (3D: one lane per wrap-group member; 2D: one thread per group) against
    the oracle, which the compiled reference pins for odd L as well: three sweeps, per
    system its own temperature, a field on the even systems of the counter runs."""
    say("odd L", "kind", kind, "L", L, "nsys", nsys, "mode", mode)
    seed, nsw = 101 + L + kind, 3
    lad = sa.Ladder(kind, L, nsys)
    lad.set_rng(mode)
    lad.seed_like_main(seed)
    Ts = [0.5 + 0.6 * j for j in range(nsys)]
    hs = [0.3 if (mode == 1 and j % 2 == 0) else 0.0 for j in range(nsys)]
    for j in range(nsys):
        lad.set_temperature(j, Ts[j], hs[j])
    jx, jy, jz, s0, g = orc.init_system(kind, L, seed)
    N = s0.size
    ref = [s0.copy() for _ in range(nsys)]
    for t in range(nsw):
        w = orc.counter_words(kind, L, seed, t) if mode == 1 else g.fill(N)
        for j in range(nsys):
            orc.sweep(kind, L, ref[j], jx, jy, jz, orc.boltz(kind, Ts[j], hs[j]), w)
    lad.sweep(nsw)
    for j in range(nsys):
        assert (lad.spins(j) == ref[j]).all(), j
    E, M = lad.measure()
    for j in range(nsys):
        e, m = orc.measure(kind, L, ref[j], jx, jy, jz)
        assert (E[j] == e).all() and (M[j] == m).all()
    lad.close()


def test_odd_L_pt_matches_oracle(sa, orc):
    """Parallel tempering on an odd lattice (replay stream) against the oracle's PT."""
    kind, L, T = 2, 9, np.array([0.5, 0.9, 1.6, 3.0])
    got = sa.pt_run(kind, L, T, 0.0, 55, 4, 6, sa.RNG_REPLAY)
    want = orc.pt_run(kind, L, T, 0.0, 55, 4, 6, rng_mode=0)
    for k in ("pairs", "us", "masks", "E", "spins"):
        assert (got[k] == want[k]).all(), k
