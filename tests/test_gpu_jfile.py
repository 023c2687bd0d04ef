"""J files (src/sys_file.cpp:10-110) through the C ABI, against files the compiled
reference's own save_J wrote (tests/golden/jfile_k*_L*.J, make_golden.py --jfiles):
a ladder seeded like main writes the reference's bytes, loading a reference file
gives the reference's couplings, and every refusal leaves the couplings unchanged."""
import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [(2, 6, 1234), (1, 8, 77), (0, 10, 5)]


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


def ref_file(kind, L):
    return os.path.join(GOLDEN, f"jfile_k{kind}_L{L}.J")


@pytest.mark.parametrize("kind,L,seed", CASES)
def test_save_J_writes_the_reference_bytes(sa, kind, L, seed):
    print("[case] save_J kind", kind, "L", L, flush=True)
    lad = sa.Ladder(kind, L, 1)
    lad.seed(seed)
    lad.init_couplings()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "J")
        lad.save_J(p)
        assert open(p, "rb").read() == open(ref_file(kind, L), "rb").read()
    lad.close()


@pytest.mark.parametrize("kind,L,seed", CASES)
def test_load_J_gives_the_reference_couplings(sa, kind, L, seed):
    print("[case] load_J kind", kind, "L", L, flush=True)
    want = sa.Ladder(kind, L, 1)
    want.seed(seed)
    want.init_couplings()
    lad = sa.Ladder(kind, L, 2)
    lad.seed(seed + 1)
    lad.init_couplings()
    lad.load_J(ref_file(kind, L))
    for a, b in zip(lad.couplings(), want.couplings()):
        assert (a == b).all()
    with tempfile.TemporaryDirectory() as d:  # and back out, byte for byte
        p = os.path.join(d, "J")
        lad.save_J(p)
        assert open(p, "rb").read() == open(ref_file(kind, L), "rb").read()
    lad.close()
    want.close()


def test_load_J_refusals_leave_couplings_unchanged(sa):
    print("[case] load_J refusals", flush=True)
    lad = sa.Ladder(2, 6, 1)
    lad.seed(9)
    lad.init_couplings()
    before = lad.couplings()
    raw = open(ref_file(2, 6), "rb").read()
    with tempfile.TemporaryDirectory() as d:
        bad = {}
        bad["missing"] = os.path.join(d, "none")
        bad["2d file"] = ref_file(1, 8)  # d = 2 into a 3D ladder (reference error -2)
        p = os.path.join(d, "N")
        with open(p, "wb") as f:  # N of another lattice (reference error -3)
            f.write(raw[:4] + np.array([125], dtype="<i8").tobytes() + raw[12:])
        bad["N"] = p
        p = os.path.join(d, "short")
        with open(p, "wb") as f:
            f.write(raw[:-8])
        bad["short"] = p
        p = os.path.join(d, "copies")
        b = bytearray(raw)
        b[12 + 8] ^= 1  # Jx[1] (site 0's +x copy) no longer equals site 1's -x bond
        with open(p, "wb") as f:
            f.write(bytes(b))
        bad["copies"] = p
        for what, path in bad.items():
            with pytest.raises(sa.SAError):
                lad.load_J(path)
            for a, c in zip(lad.couplings(), before):
                assert (a == c).all(), what
    lad.close()
