"""Text output of the device PT driver against the reference's own files.

parallel_tempering with a stride prints a line of 64 energies per temperature
every `stride` sweeps into files named from a pattern, the E_min file and the
acceptance report (src/parallel_tempering.hpp:117-139,164-170,203-262).  With the
replay RNG the device driver must write the same bytes (tests/golden/pt_files.json,
written by the compiled reference)."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def runs():
    with open(os.path.join(HERE, "pt_files.json")) as f:
        return json.load(f)["runs"]


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


@pytest.mark.parametrize("c", runs(), ids=lambda c: f"k{c['kind']}_L{c['L']}_stride{c['stride']}")
def test_pt_files_match_reference(sa, c, tmp_path):
    print("[case] pt files", c["kind"], c["L"], c["stride"], flush=True)
    pt = sa.PT(c["kind"], c["L"], c["T"], c["h"], c["seed"], c["sweeps"], sa.RNG_REPLAY)
    pt.set_output(tmp_path / "pt_%.dat")
    pt.step_stride(c["nsteps"], c["stride"])
    pt.print_min()
    report = pt.swap_report()
    pt.close_output()
    assert report == c["swap_text"]
    got = {f: (tmp_path / f).read_text() for f in sorted(os.listdir(tmp_path))}
    assert sorted(got) == sorted(c["files"])
    for name, text in c["files"].items():
        assert got[name] == text, name
    pt.close()
