"""Checkpoints (.tempsys) of the device PT driver against the reference's own.

With the replay RNG the device driver's parallel_tempering::save must write the
reference's checkpoint byte for byte (tests/golden/ckpt_*.tempsys, written by the
compiled reference after the same steps), and loading a reference checkpoint must
continue the run exactly as the reference did (pairs, uniforms, masks, spins).
With the counter RNG a save/load round trip continues identically."""
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import f64, words

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


def ckpts():
    with open(os.path.join(HERE, "ckpt_vectors.json")) as f:
        return json.load(f)["checkpoints"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


@pytest.mark.parametrize("c", ckpts(), ids=lambda c: c["file"])
def test_save_writes_the_reference_checkpoint(sa, c, tmp_path):
    say("save", c["file"])
    pt = sa.PT(c["kind"], c["L"], c["T"], c["h"], c["seed"], c["sweeps"], sa.RNG_REPLAY)
    pt.step(c["nsteps1"])
    out = tmp_path / "dev.tempsys"
    pt.save(out)
    with open(os.path.join(HERE, c["file"]), "rb") as f:
        want = f.read()
    got = out.read_bytes()
    assert len(got) == len(want)
    assert got == want
    pt.close()


@pytest.mark.parametrize("c", ckpts(), ids=lambda c: c["file"])
def test_load_continues_the_reference_run(sa, c):
    say("load", c["file"])
    # a driver built with other seeds: everything comes from the checkpoint
    pt = sa.PT(c["kind"], c["L"], c["T"], c["h"], 99, c["sweeps"], sa.RNG_REPLAY)
    pt.load(os.path.join(HERE, c["file"]))
    pt.step(c["nsteps2"])
    r = pt.result()
    assert [int(x) for x in r["pairs"]] == c["pairs"]
    assert (r["us"] == f64(c["us"])).all()
    assert (r["masks"] == words(c["masks"])).all()
    assert [sha(s) for s in r["spins"]] == c["spins_sha"]
    pt.close()


def test_counter_rng_round_trip(sa, tmp_path):
    T = [0.5, 0.8, 1.3, 2.0]
    a = sa.PT(2, 8, T, 0.0, 77, 5, sa.RNG_COUNTER)
    a.step(4)
    a.save(tmp_path / "c.tempsys")
    a.step(6)
    ra = a.result()
    b = sa.PT(2, 8, T, 0.0, 1, 5, sa.RNG_COUNTER)
    b.load(tmp_path / "c.tempsys")
    b.step(6)
    rb = b.result()
    assert (ra["pairs"][4:] == rb["pairs"]).all() and (ra["masks"][4:] == rb["masks"]).all()
    assert (ra["spins"] == rb["spins"]).all()
    a.close()
    b.close()


def test_load_after_update_T_into_one_ladder(sa, tmp_path):
    """A checkpoint written after update_T holds inserted systems whose engines
    restarted at seed 1234 (the reference's clone): loading it into a driver whose
    systems are slots of one ladder must not overwrite the shared engine with theirs
    -- those systems move onto ladders of their own -- and the run continues as the
    original does (replay stream, compared step by step)."""
    say("load after update_T")
    T = [0.4, 1.2, 3.0]
    a = sa.PT(2, 8, T, 0.0, 77, 5, sa.RNG_REPLAY)
    a.step(20)
    n = a.update_temperatures()
    assert n > len(T)
    a.step(3)
    a.save(tmp_path / "u.tempsys")
    a.step(6)
    ra = a.result()
    b = sa.PT(2, 8, a.temperatures(), 0.0, 1, 5, sa.RNG_REPLAY)
    b.load(tmp_path / "u.tempsys")
    b.step(6)
    rb = b.result()
    assert (ra["pairs"][-6:] == rb["pairs"]).all()
    assert (ra["us"][-6:] == rb["us"]).all()
    assert (ra["masks"][-6:] == rb["masks"]).all()
    assert (ra["spins"] == rb["spins"]).all()
    assert (ra["E"] == rb["E"]).all()
    a.close()
    b.close()
