"""The ground-state loop fixtures (tests/golden/pt_ground.json, written by the
compiled reference: make_golden.py --ground) checked against the rule that
parallel_tempering::update_T states (src/parallel_tempering.hpp:384-413): every
neighbour pair whose exchange rate is below 0.3 gets the temperature
(T_i + T_i+1) / 2 inserted between them, each old pair checked once.  This pins the
fixture the GPU test (tests/test_gpu_pt_ground.py) compares against."""
import json
import os
import struct

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def runs():
    with open(os.path.join(HERE, "pt_ground.json")) as f:
        return json.load(f)["runs"]


def f64(h):
    """a double from the hex of its bit pattern"""
    return struct.unpack("<d", struct.pack("<Q", int(h, 16)))[0]


def insert_rule(T, acc):
    out = [T[0]]
    for i in range(len(T) - 1):
        if acc[i] < 0.3:
            out.append((T[i + 1] + T[i]) / 2.0)
        out.append(T[i + 1])
    return out


@pytest.mark.parametrize("r", runs(), ids=lambda r: f"k{r['kind']}_L{r['L']}")
def test_update_T_rule_on_reference_fixture(r):
    T = sorted(r["T"])
    assert len(r["acc"]) == r["rounds"] == len(r["T_rounds"]) == len(r["nT"])
    for q in range(r["rounds"]):
        acc = [f64(h) for h in r["acc"][q]]
        assert len(acc) == len(T) - 1
        assert all(0.0 <= a <= 1.0 for a in acc)
        want = insert_rule(T, acc)
        got = [f64(h) for h in r["T_rounds"][q]]
        assert got == want and r["nT"][q] == len(want)
        assert len(want) > len(T) or all(a >= 0.3 for a in acc)
        T = want
    assert len(r["spins_sha"]) == len(T) == len(r["E"])
