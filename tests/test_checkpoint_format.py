"""The .tempsys checkpoint format (parallel_tempering::save + cpu_*::save records,
src/parallel_tempering.hpp:280-315, src/cpu_3d.cpp:449-461, src/cpu_2dp.cpp:444-455).

CPU side: the reference's own checkpoints (tests/golden/ckpt_*.tempsys, written by
the compiled reference, see make_golden.py) are parsed field by field and their
contents checked against the oracle: spins after the saved steps, couplings in the
duplicated layout, the engine state, the sorted temperatures.  This pins the layout
the device driver's sa_pt_save / sa_pt_load follow (tests/test_gpu_checkpoint.py).
"""
import json
import os
import struct

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ckpts():
    with open(os.path.join(HERE, "ckpt_vectors.json")) as f:
        return json.load(f)["checkpoints"]


class Reader:
    def __init__(self, b: bytes):
        self.b, self.o = b, 0

    def take(self, n):
        v = self.b[self.o:self.o + n]
        assert len(v) == n, "truncated record"
        self.o += n
        return v

    def fmt(self, f):
        return struct.unpack("<" + f, self.take(struct.calcsize("<" + f)))

    def arr(self, dtype, n):
        return np.frombuffer(self.take(np.dtype(dtype).itemsize * n), dtype=dtype)


def parse(b: bytes, kind: int):
    """Fields of a parallel_tempering::save stream."""
    r = Reader(b)
    out = {}
    out["steps"], = r.fmt("q")
    out["print_E"], out["print_min"] = r.fmt("??")
    out["gen32"] = r.arr("<u8", 624)  # std::mt19937 keeps uint_fast32_t words: 8 bytes on LP64
    out["gen32_pos"], = r.fmt("Q")
    out["uni_int"] = r.fmt("ii")
    out["uni_double"] = r.fmt("dd")
    n, = r.fmt("Q")
    out["fname"] = r.take(n).decode()
    n, = r.fmt("Q")
    out["T"] = r.arr("<f8", n)
    n, = r.fmt("Q")
    out["E_min"] = r.arr("<f4", n)
    n, = r.fmt("Q")
    out["num_switch"] = [r.fmt("dq") for _ in range(n)]
    nE, = r.fmt("Q")
    d = 3 if kind == 2 else 2
    out["sys"] = []
    for _ in range(nE):
        m, = r.fmt("Q")
        E = r.arr("<f4", m)
        L, = r.fmt("i")
        T, h = r.fmt("dd")
        N = L ** (3 if kind == 2 else 2)
        s = r.arr("<u8", N)
        J = [r.arr("<u8", 2 * N) for _ in range(d)]
        gen = r.arr("<u8", 312)
        pos, = r.fmt("Q")
        ud = r.fmt("dd")
        ui = r.fmt("ii")
        out["sys"].append(dict(E=E, L=L, T=T, h=h, s=s, J=J, gen=gen, pos=pos, ud=ud, ui=ui))
    assert r.o == len(b), "trailing bytes"
    return out


@pytest.mark.parametrize("c", ckpts(), ids=lambda c: c["file"])
def test_reference_checkpoint_layout(orc, c):
    kind, L = c["kind"], c["L"]
    with open(os.path.join(HERE, c["file"]), "rb") as f:
        p = parse(f.read(), kind)
    nT = len(c["T"])
    N = L ** (3 if kind == 2 else 2)
    assert p["steps"] == c["sweeps"] and p["fname"] == c["fname"]
    assert (p["T"] == np.sort(np.array(c["T"]))).all()
    assert p["uni_int"] == (0, nT - 2) and p["uni_double"] == (0.0, 1.0)
    assert len(p["num_switch"]) == nT - 1 and len(p["sys"]) == nT
    assert sum(n for _, n in p["num_switch"]) == c["nsteps1"]  # one exchange attempt per step
    # the oracle's PT run of the same steps gives the saved spins
    ref = orc.pt_run(kind, L, np.array(c["T"]), c["h"], c["seed"], c["sweeps"], c["nsteps1"])
    jx, jy, jz, s0, g = orc.init_system(kind, L, c["seed"])
    for j, sy in enumerate(p["sys"]):
        assert sy["L"] == L and sy["T"] == p["T"][j] and sy["h"] == c["h"]
        assert (sy["s"] == ref["spins"][j]).all()
        # duplicated coupling layout: entry 2i is the deduplicated bond of site i
        assert (sy["J"][0][0::2] == jx).all() and (sy["J"][1][0::2] == jy).all()
        if kind == 2:
            assert (sy["J"][2][0::2] == jz).all()
        assert sy["ud"] == (0.0, 1.0) and sy["ui"] == (0, N - 1)
        # energies recorded for the system are the oracle's (measure() values)
        assert (sy["E"] == ref["E"][j]).all() or j in (int(ref["pairs"][-1]), int(ref["pairs"][-1]) + 1)
    # every system's engine sits at the same stream position: N (init) + sweeps x N words
    st = [tuple(sy["gen"][:4]) + (sy["pos"],) for sy in p["sys"]]
    assert len(set(st)) == 1
    eng = orc.Mt64(c["seed"])
    eng.fill(N + c["nsteps1"] * c["sweeps"] * N)
    assert (np.frombuffer(bytes(eng.st), dtype="<u8")[:312] == p["sys"][0]["gen"]).all()
