"""The pieces of main_cpu's ground-state loop (-r, src/main_cpu.cpp:344-434) on the
device driver against the compiled reference (tests/golden/pt_ground.json):
rounds of PT steps, get_E_min, update_T (temperatures inserted where the exchange
rate is below 0.3, src/parallel_tempering.hpp:384-413: the new system is a copy of
its hotter neighbour on a ladder of its own) and reset (new spins from every
system's engine, src/parallel_tempering.hpp:265-277), then steps once more.  Replay
stream; bar: bit-exact rates, temperatures, E_min, final spins (SHA-256) and
energies of every system."""
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import f32, f64

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def runs():
    with open(os.path.join(HERE, "pt_ground.json")) as f:
        return json.load(f)["runs"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


@pytest.mark.parametrize("r", runs(), ids=lambda r: f"k{r['kind']}_L{r['L']}")
def test_ground_loop_matches_reference(sa, r):
    print("[case] ground loop kind", r["kind"], "L", r["L"], "T", r["T"], flush=True)
    pt = sa.PT(r["kind"], r["L"], r["T"], r["h"], r["seed"], r["sweeps"], sa.RNG_REPLAY)
    for q in range(r["rounds"]):
        pt.step(r["nsteps"])
        assert (pt.E_min() == f32(r["E_min"][q])).all(), q
        assert (pt.acceptance() == f64(r["acc"][q])).all(), q
        assert pt.update_temperatures() == r["nT"][q]
        assert (pt.temperatures() == f64(r["T_rounds"][q])).all(), q
        pt.reset()
        assert (pt.E_min() == 0).all()
    pt.step(r["nsteps"])
    res = pt.result()
    assert res["spins"].shape[0] == len(r["spins_sha"])
    for j, want in enumerate(r["spins_sha"]):
        assert sha(res["spins"][j]) == want, j
    for j, e in enumerate(r["E"]):
        assert (res["E"][j] == f32(e)).all(), j
    pt.close()
