"""The command-line driver (simulated_annealing_amd/sa_main, the reference's
annealing_gpu flags, src/main_gpu.cu:56-209): a geometric anneal prints, after its
info header, what the annealing driver prints for the same run (sa_anneal_geometric,
itself byte-identical to the compiled reference); -J / -j move the couplings through
a J file without changing the run; -t reports seconds per replica spin flip."""
import os
import subprocess
import tempfile

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "simulated_annealing_amd", "sa_main")


def run(args, cwd):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, cwd=cwd)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("dim,kind,L", [("3", 2, 6), ("2p", 0, 12), ("2", 1, 16)])
def test_cli_anneal_prints_the_driver_lines(dim, kind, L):
    import simulated_annealing_amd as sa

    print("[case] sa_main anneal -d", dim, "-L", L, flush=True)
    with tempfile.TemporaryDirectory() as d:
        args = ["-d", dim, "-L", str(L), "-T", "3.0", "-a", "0.9", "-S", "4", "-s", "6", "--seed", "1234"]
        out = run(args, d)
        assert out.startswith("# ")
        body = out.split("\n", 2)[2]  # after the command line and version lines
        want = sa.anneal_geometric(kind, L, 3.0, 0.9, 4, 6, 1, 0.0, 1234, sa.RNG_REPLAY, precision=0)
        assert body == want
        jf = os.path.join(d, "J.bin")
        out2 = run(args + ["-J", jf], d)
        assert out2.split("\n", 2)[2] == want and os.path.getsize(jf) > 12
        out3 = run(args + ["-j", jf, "--seed", "1234"], d)
        assert out3.split("\n", 2)[2] == want


def test_cli_timing_report():
    print("[case] sa_main -t", flush=True)
    with tempfile.TemporaryDirectory() as d:
        out = run(["-d", "3", "-L", "32", "-s", "20", "-t"], d).strip().splitlines()
    assert out[2].split("\t") == ["3", "32", "20"]
    mean, std, tmin, tmax = (float(v) for v in out[3].split("\t"))
    assert 0 < tmin <= mean <= tmax and std >= 0


def test_cli_refuses_unknown_and_unprovided_flags():
    for args in (["-x"], ["-r", "3", "-T", "1"], ["-T", "1", "-P", "1", "-P", "2"], ["-P", "1"]):
        r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode == 2 and "sa_main:" in r.stderr, args
