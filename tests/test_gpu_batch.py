"""Batched instances (sa_batch_*) against each instance run alone.

Every instance of a batch must end with the spin words of its own ladder swept
alone through the plain path (sa_ladder_set_temperature + sa_ladder_sweep per
sweep, counter RNG), whatever the mix, the block size, the graph chunk length or
the way the run is split into calls.  A small mix is also checked against the
oracle directly.  Bar: bit-exact spin words and integer energies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


def lone(sa, kind, L, nsys, seed, T, per_sweep, nsweeps, h=0.0):
    """The instance alone through the plain path; returns (spins [nsys, N], E, M)."""
    lad = sa.Ladder(kind, L, nsys)
    lad.set_rng(sa.RNG_COUNTER)
    lad.seed_like_main(seed)
    T = np.asarray(T, dtype=np.float64).reshape(-1, nsys) if per_sweep else np.asarray(T).reshape(1, nsys)
    if not per_sweep:
        for j in range(nsys):
            lad.set_temperature(j, T[0, j], h)
        lad.sweep(nsweeps)
    else:
        for t in range(nsweeps):
            for j in range(nsys):
                lad.set_temperature(j, T[t, j], h)
            lad.sweep(1)
    spins = np.stack([lad.spins(j) for j in range(nsys)])
    E, M = lad.measure()
    lad.close()
    return spins, E, M


# (kind, L, nsys, per_sweep, nsweeps, h)
MIX = [
    (1, 8, 1, True, 17, 0.0),
    (1, 12, 1, True, 5, 0.0),
    (1, 32, 1, True, 40, 0.0),
    (1, 48, 2, False, 23, 0.0),
    (0, 10, 1, True, 31, 0.0),
    (0, 16, 1, False, 12, 0.3),
    (2, 4, 1, True, 9, 0.0),
    (2, 6, 3, False, 26, 0.0),
    (2, 16, 1, True, 40, 0.0),
    (2, 24, 1, True, 19, 0.0),
    (2, 8, 1, False, 33, -0.4),
    (1, 20, 1, True, 0, 0.0),
]


def build_mix(sa, b, mix, seed0):
    rng = np.random.default_rng(seed0)
    out = []
    for q, (kind, L, nsys, per_sweep, nsw, h) in enumerate(mix):
        seed = seed0 + 17 * q
        if per_sweep:
            # per system its own linear anneal (a different end point each)
            T = np.stack([sa.linear_schedule(3.0, 0.1 + 0.2 * j, nsw) for j in range(nsys)], axis=1)
            i = b.add(kind, L, seed, nsw, schedule=T, nsys=nsys, h=h)
        else:
            T = rng.uniform(0.4, 2.5, size=nsys)
            i = b.add(kind, L, seed, nsw, T=T, nsys=nsys, h=h)
        out.append((i, kind, L, nsys, per_sweep, nsw, h, T, seed))
    return out


def check_mix(sa, b, insts):
    for (i, kind, L, nsys, per_sweep, nsw, h, T, seed) in insts:
        say("batch vs lone", "kind", kind, "L", L, "nsys", nsys, "sweeps", nsw, "h", h)
        assert b.sweeps_done(i) == nsw
        ref_s, ref_E, ref_M = lone(sa, kind, L, nsys, seed, T, per_sweep, nsw, h)
        for j in range(nsys):
            assert (b.spins(i, j) == ref_s[j]).all(), (kind, L, j)
        E, M = b.measure(i)
        assert (E == ref_E).all() and (M == ref_M).all()


@pytest.mark.parametrize("block,chunk", [(0, 0), (64, 8), (256, 3), (128, 32)])
def test_batch_matches_lone_instances(sa, block, chunk):
    say("batch mix", "block", block, "chunk", chunk)
    b = sa.Batch()
    if block:
        b.tune(sa.TUNE_BATCH_BLOCK, block)
    if chunk:
        b.tune(sa.TUNE_BATCH_CHUNK, chunk)
    insts = build_mix(sa, b, MIX, 4000 + block + chunk)
    assert b.run() == 0
    check_mix(sa, b, insts)
    b.close()


def test_batch_split_runs_equal_one_run(sa):
    say("batch split runs")
    b = sa.Batch()
    insts = build_mix(sa, b, MIX, 777)
    left = b.run(7)
    assert left > 0
    for (i, *_rest) in insts:
        assert b.sweeps_done(i) == min(7, _rest[4])
    b.run(11)
    assert b.run() == 0
    check_mix(sa, b, insts)
    b.close()


def test_batch_add_and_remove_between_runs(sa):
    say("batch add/remove")
    b = sa.Batch()
    first = build_mix(sa, b, MIX[:6], 99)
    b.run(10)
    b.remove(first[2][0])
    b.tune(sa.TUNE_BATCH_CHUNK, 4)
    late = build_mix(sa, b, MIX[6:], 1234)
    assert b.run() == 0
    check_mix(sa, b, [x for x in first if x[0] != first[2][0]] + late)
    with pytest.raises(sa.SAError):
        b.sweeps_done(first[2][0])
    b.close()


def test_batch_matches_oracle(sa, orc):
    """Two instances checked against the oracle's counter-word sweeps directly."""
    b = sa.Batch()
    cases = [(2, 6, 11, 21), (1, 8, 13, 22), (0, 12, 9, 23)]
    ids = []
    for kind, L, nsw, seed in cases:
        ids.append(b.add(kind, L, seed, nsw, schedule=sa.linear_schedule(3.0, 0.1, nsw)))
    assert b.run() == 0
    for (kind, L, nsw, seed), i in zip(cases, ids):
        say("batch vs oracle", "kind", kind, "L", L, "sweeps", nsw)
        jx, jy, jz, s, _ = orc.init_system(kind, L, seed)
        for t, T in enumerate(sa.linear_schedule(3.0, 0.1, nsw)):
            orc.sweep(kind, L, s, jx, jy, jz, orc.boltz(kind, T), orc.counter_words(kind, L, seed, t))
        assert (b.spins(i) == s).all()
    b.close()


def test_batch_rejects_bad_instances(sa):
    b = sa.Batch()
    with pytest.raises(sa.SAError):
        b.add(2, 2, 1, 3, T=np.ones(1))  # L < 3 is not supported on the device
    with pytest.raises(sa.SAError):
        b.tune(sa.TUNE_BATCH_BLOCK, 96)
    with pytest.raises(sa.SAError):
        b.tune(sa.TUNE_BATCH_FUSED, 2)  # the fused batch sweep is switched off
    b.tune(sa.TUNE_BATCH_FUSED, 1)
    b.close()


# odd L: the wrap groups walked inside the batch kernel
ODD_MIX = [
    (2, 9, 1, True, 11, 0.0),
    (2, 5, 2, False, 14, 0.0),
    (2, 3, 1, True, 6, 0.0),
    (1, 15, 1, True, 17, 0.0),
    (1, 33, 2, False, 9, 0.0),
    (0, 13, 1, True, 12, 0.0),
    (0, 11, 1, False, 10, 0.25),
    (2, 7, 1, False, 8, -0.3),
    (2, 8, 1, True, 10, 0.0),   # even instances in the same launches
    (1, 16, 1, True, 13, 0.0),
]


def test_odd_L_instances_match_oracle(sa, orc):
    """Odd-L instances of every kind, beside even ones, against the oracle's
    counter-word sweeps."""
    b = sa.Batch()
    cases = [(2, 9, 7, 61), (2, 5, 4, 62), (1, 15, 9, 63), (1, 3, 5, 64), (0, 13, 6, 65), (2, 6, 5, 66),
             (2, 37, 3, 67)]
    ids = [b.add(kind, L, seed, nsw, schedule=sa.linear_schedule(3.0, 0.1, nsw)) for kind, L, nsw, seed in cases]
    assert b.run() == 0
    for (kind, L, nsw, seed), i in zip(cases, ids):
        say("odd-L batch vs oracle", "kind", kind, "L", L, "sweeps", nsw)
        jx, jy, jz, s, _ = orc.init_system(kind, L, seed)
        for t, T in enumerate(sa.linear_schedule(3.0, 0.1, nsw)):
            orc.sweep(kind, L, s, jx, jy, jz, orc.boltz(kind, T), orc.counter_words(kind, L, seed, t))
        assert (b.spins(i) == s).all()
    b.close()


def test_odd_L_mix_matches_plain_path(sa):
    """Odd and even instances (fixed and per-sweep temperatures, fields) run in chunks
    give the words and energies of each instance swept alone."""
    say("odd-L mix")
    b = sa.Batch()
    insts = build_mix(sa, b, ODD_MIX, 77)
    b.tune(sa.TUNE_BATCH_CHUNK, 3)
    assert b.run() == 0
    check_mix(sa, b, insts)
    b.close()


def test_odd_L_pt_matches_pt_driver(sa):
    """Parallel tempering on odd lattices inside the batch: swap sequence and spins of
    the PT driver run alone."""
    b = sa.Batch()
    pts = []
    for q, (kind, L, nT, every, nsteps, h) in enumerate([(2, 7, 4, 5, 4, 0.0), (0, 11, 3, 6, 3, 0.2),
                                                         (1, 9, 5, 4, 5, 0.0)]):
        T = 0.5 * 4.0 ** (np.arange(nT)[::-1] / (nT - 1))
        seed = 700 + q
        pts.append((b.add_pt(kind, L, T, seed, every, nsteps, h=h), kind, L, nT, every, nsteps, h, T, seed))
    assert b.run() == 0
    for (i, kind, L, nT, every, nsteps, h, T, seed) in pts:
        say("odd-L batch PT vs driver", "kind", kind, "L", L, "nT", nT)
        ref = sa.PT(kind, L, T, h, seed, every, sa.RNG_COUNTER)
        ref.step(nsteps)
        r = ref.result()
        got = b.pt_result(i)
        assert (got["pairs"] == r["pairs"]).all()
        assert (got["us"] == r["us"]).all()
        assert (got["masks"] == r["masks"]).all()
        for j in range(nT):
            assert (b.spins(i, j) == r["spins"][j]).all(), j
        ref.close()
    b.close()


def test_uneven_mix_of_64_instances(sa):
    """A scaled-down form of the 512-instance workload: 60% 2D, 40% 3D, linear anneals."""
    rng = np.random.default_rng(5)
    b = sa.Batch()
    insts = []
    for q in range(64):
        if rng.random() < 0.6:
            kind, L = 1, int(rng.choice([32, 48, 64, 96, 128]))
        else:
            kind, L = 2, int(rng.choice([16, 24, 32, 48]))
        nsw = int(rng.choice([5, 10, 20]))
        T = sa.linear_schedule(3.0, 0.1, nsw)
        i = b.add(kind, L, 10_000 + q, nsw, schedule=T)
        insts.append((i, kind, L, 1, True, nsw, 0.0, T, 10_000 + q))
    say("64-instance mix")
    assert b.run() == 0
    check_mix(sa, b, insts)
    b.close()


# (kind, L, nT, every, nsteps, h)
PT_MIX = [
    (2, 8, 4, 5, 6, 0.0),
    (1, 16, 6, 10, 4, 0.0),
    (0, 12, 3, 7, 5, 0.2),
    (2, 6, 8, 10, 3, -0.3),
    (1, 32, 5, 3, 9, 0.0),
]


def test_batch_pt_matches_pt_driver(sa):
    """Parallel-tempering instances mixed with anneals: each has the swap sequence
    (pairs, uniforms, lane masks) and final spins of the PT driver run alone."""
    b = sa.Batch()
    anneals = build_mix(sa, b, MIX[:5], 31)
    pts = []
    for q, (kind, L, nT, every, nsteps, h) in enumerate(PT_MIX):
        T = 0.5 * 4.0 ** (np.arange(nT)[::-1] / (nT - 1))  # descending: the driver sorts
        seed = 500 + q
        i = b.add_pt(kind, L, T, seed, every, nsteps, h=h)
        pts.append((i, kind, L, nT, every, nsteps, h, T, seed))
    assert b.run() == 0
    check_mix(sa, b, anneals)
    for (i, kind, L, nT, every, nsteps, h, T, seed) in pts:
        say("batch PT vs driver", "kind", kind, "L", L, "nT", nT, "every", every, "steps", nsteps, "h", h)
        assert b.sweeps_done(i) == every * nsteps
        ref = sa.PT(kind, L, T, h, seed, every, sa.RNG_COUNTER)
        ref.step(nsteps)
        r = ref.result()
        got = b.pt_result(i)
        assert (got["pairs"] == r["pairs"]).all()
        assert (got["us"] == r["us"]).all()
        assert (got["masks"] == r["masks"]).all()
        for j in range(nT):
            assert (b.spins(i, j) == r["spins"][j]).all(), j
        # energies now (the driver's E keeps the pre-exchange values of the last pair,
        # as the reference's E_ does): measure the driver's ladder directly
        E, M = b.measure(i)
        E_ref = np.zeros(nT * 64, dtype=np.int32)
        M_ref = np.zeros(nT * 64, dtype=np.int32)
        sa.check(sa.lib().sa_ladder_measure(ref.ladder_handle(), E_ref, M_ref), "measure")
        assert (E.reshape(-1) == E_ref).all() and (M.reshape(-1) == M_ref).all()
        # systems outside the last exchanged pair agree with the driver's E as well
        last = int(got["pairs"][-1])
        keep = [j for j in range(nT) if j not in (last, last + 1)]
        assert ((E[keep] + h * M[keep]).astype(np.float32) == r["E"][keep]).all()
        ref.close()
    b.close()


# 3D instances that take the fused sweep (L % 4 == 0, h = 0) beside ones that cannot
FUSED_MIX = [
    (2, 4, 1, True, 9, 0.0),
    (2, 8, 2, False, 13, 0.0),
    (2, 12, 1, True, 17, 0.0),
    (2, 16, 3, False, 11, 0.0),
    (2, 24, 1, True, 21, 0.0),
    (2, 32, 2, True, 7, 0.0),
    (2, 48, 1, True, 6, 0.0),
    (2, 6, 1, True, 10, 0.0),
    (2, 16, 1, True, 12, 0.3),
    (1, 32, 1, True, 15, 0.0),
]


FUSED_OFF = pytest.mark.skip(reason="the fused batch sweep is switched off: one run of its test stopped "
                              "making progress and no cause was found; it stays off the GPU until one is")


@FUSED_OFF
@pytest.mark.parametrize("fused,chunk", [(2, 1), (2, 5), (2, 32), (1, 8)])
def test_batch_fused_sweep_matches_lone(sa, fused, chunk):
    say("batch fused", fused, "chunk", chunk)
    b = sa.Batch()
    b.tune(sa.TUNE_BATCH_FUSED, fused)
    b.tune(sa.TUNE_BATCH_CHUNK, chunk)
    b.tune(sa.TUNE_BATCH_BLOCK, 128)
    insts = build_mix(sa, b, FUSED_MIX, 900 + chunk)
    assert b.run() == 0
    check_mix(sa, b, insts)
    b.close()


@FUSED_OFF
def test_batch_fused_mode_switches_mid_run(sa):
    """Odd and even numbers of fused sweeps between switches: the current spin array
    is tracked across fused (ping-pong) and half-sweep (in place) chunks."""
    say("batch fused switching")
    b = sa.Batch()
    b.tune(sa.TUNE_BATCH_BLOCK, 64)
    insts = build_mix(sa, b, FUSED_MIX, 321)
    for fused, chunk, n in [(2, 3, 3), (1, 2, 4), (2, 1, 1), (2, 4, 5), (1, 3, 3)]:
        b.tune(sa.TUNE_BATCH_FUSED, fused)
        b.tune(sa.TUNE_BATCH_CHUNK, chunk)
        b.run(n)
    b.tune(sa.TUNE_BATCH_FUSED, 2)
    assert b.run() == 0
    check_mix(sa, b, insts)
    b.close()


@FUSED_OFF
def test_batch_fused_pt_matches_pt_driver(sa):
    """Parallel tempering through the fused sweep (even nT: two systems per workgroup)."""
    b = sa.Batch()
    b.tune(sa.TUNE_BATCH_FUSED, 2)
    cases = [(2, 16, 6, 5, 4), (2, 8, 3, 4, 5), (2, 24, 4, 10, 2)]
    ids = []
    for q, (kind, L, nT, every, nsteps) in enumerate(cases):
        T = 0.5 * 4.0 ** (np.arange(nT) / (nT - 1))
        ids.append((b.add_pt(kind, L, T, 60 + q, every, nsteps), kind, L, nT, every, nsteps, T, 60 + q))
    assert b.run() == 0
    for (i, kind, L, nT, every, nsteps, T, seed) in ids:
        say("fused PT", L, nT)
        ref = sa.PT(kind, L, T, 0.0, seed, every, sa.RNG_COUNTER)
        ref.step(nsteps)
        r = ref.result()
        got = b.pt_result(i)
        assert (got["pairs"] == r["pairs"]).all() and (got["masks"] == r["masks"]).all()
        for j in range(nT):
            assert (b.spins(i, j) == r["spins"][j]).all(), j
        ref.close()
    b.close()
