"""The C++ drop-in (tests/dropin/pt_dropin.cpp: the reference's main_gpu set-up with
one gpu_3d / gpu_2d per temperature under parallel_tempering<gpu_sys>, plus a J-file
save/load round trip) against the compiled reference's parallel-tempering fixtures:
pairs, uniforms, masks, energies and spins, bit for bit (replay RNG, the default of
the drop-in classes)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import dropin_build
from helpers import f32, f64, sha, words

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [0, 1, 3])  # 3D L=6 (4 T), 3D L=8 (8 T), 2p L=12 (6 T)
def test_dropin_pt_matches_reference(golden, case):
    c = golden["pt"][case]
    assert c["kind"] in (0, 2) and c["h"] == 0.0
    print("[case] drop-in PT kind", c["kind"], "L", c["L"], "nT", len(c["T"]), flush=True)
    with tempfile.TemporaryDirectory() as d:
        exe = dropin_build.build(os.path.join(d, "pt_dropin"))
        out = os.path.join(d, "spins.bin")
        args = [exe, str(c["kind"]), str(c["L"]), str(c["seed"]), str(c["sweeps"]), str(c["nsteps"]), out,
                *[repr(float(t)) for t in c["T"]]]
        r = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=d)
        assert r.returncode == 0, r.stderr
        assert "Error" not in r.stderr, r.stderr
        steps = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("step ")]
        E = [ln.split()[2:] for ln in r.stdout.splitlines() if ln.startswith("E ")]
        spins = np.fromfile(out, dtype="<u8").reshape(len(c["T"]), -1)
    assert [int(s[1]) for s in steps] == c["pairs"]
    assert (np.array([int(s[2], 16) for s in steps], dtype=np.uint64).view(np.float64) == f64(c["us"])).all()
    assert (np.array([int(s[3], 16) for s in steps], dtype=np.uint64) == words(c["masks"])).all()
    for j, e in enumerate(c["E"]):
        assert (f32(E[j]) == f32(e)).all(), j
    assert [sha(s) for s in spins] == c["spins_sha"]
