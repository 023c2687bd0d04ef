"""Synthetic workloads of the batched and serving forms (seeded, reproducible).

uneven_instances: independent disorder realisations of uneven size, 60% 2D
(periodic, L in {32, 48, 64, 96, 128}) and 40% 3D (L in {16, 24, 32, 48}), each a
linear anneal 3.0 -> 0.1 over a sweep count from {500, 1000, 2000}.

serving_requests: a Poisson stream of anneal / parallel-tempering requests, 2D
with L log-uniform in [16, 256] or 3D with L log-uniform in [8, 64] (rounded to
the nearest integer, odd L included), sweeps log-uniform in
[100, 10000]; a PT request has 8..32 temperatures geometric in [0.5, 2.0] and
exchanges every 10 sweeps.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

KIND_2D, KIND_3D = 1, 2


@dataclass
class Instance:
    kind: int
    L: int
    nsweeps: int
    seed: int

    @property
    def sites(self) -> int:
        return self.L ** (3 if self.kind == KIND_3D else 2)


def uneven_instances(n: int = 512, seed: int = 2024, sweeps=(500, 1000, 2000), scale: float = 1.0,
                     seed_base: int = 1000):
    """n instances; exactly round(0.6 n) are 2D.  Instance i is seeded seed_base + i."""
    rng = np.random.default_rng(seed)
    n2 = int(round(0.6 * n))
    kinds = np.array([KIND_2D] * n2 + [KIND_3D] * (n - n2))
    rng.shuffle(kinds)
    out = []
    for i, kind in enumerate(kinds):
        if kind == KIND_2D:
            L = int(rng.choice([32, 48, 64, 96, 128]))
        else:
            L = int(rng.choice([16, 24, 32, 48]))
        nsw = max(1, int(int(rng.choice(sweeps)) * scale))
        out.append(Instance(int(kind), L, nsw, seed_base + i))
    return out


def anneal_schedule(nsweeps: int, T0: float = 3.0, T1: float = 0.1) -> np.ndarray:
    """Temperature of each sweep of a linear anneal from T0 to T1."""
    if nsweeps <= 1:
        return np.full(max(nsweeps, 0), T0, dtype=np.float64)
    return T0 + (T1 - T0) * np.arange(nsweeps, dtype=np.float64) / (nsweeps - 1)


def geometric_ladder(nT: int, T0: float = 0.5, T1: float = 2.0) -> np.ndarray:
    return T0 * (T1 / T0) ** (np.arange(nT, dtype=np.float64) / (nT - 1))


@dataclass
class Request:
    arrival: float  # seconds after the start of the stream
    kind: int
    L: int
    mode: int  # 0 anneal, 1 parallel tempering
    nsweeps: int
    nT: int  # systems (1 for an anneal)
    seed: int

    @property
    def sites(self) -> int:
        return self.L ** (3 if self.kind == KIND_3D else 2)

    @property
    def attempts(self) -> int:
        return self.sites * 64 * self.nT * self.nsweeps


def _log_uniform_L(rng, lo: int, hi: int) -> int:
    return int(min(hi, max(lo, round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))))


def serving_requests(n: int = 20000, rate: float = 50.0, seed: int = 7, p3d: float = 0.5, p_pt: float = 0.5,
                     sweeps=(100, 10000), pt_every: int = 10, seed_base: int = 100000):
    """n requests with exponential inter-arrival times of mean 1/rate seconds."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        if rng.random() < p3d:
            kind, L = KIND_3D, _log_uniform_L(rng, 8, 64)
        else:
            kind, L = KIND_2D, _log_uniform_L(rng, 16, 256)
        nsw = int(round(np.exp(rng.uniform(np.log(sweeps[0]), np.log(sweeps[1])))))
        if rng.random() < p_pt:
            mode, nT = 1, int(rng.integers(8, 33))
            nsw = max(pt_every, (nsw // pt_every) * pt_every)
        else:
            mode, nT = 0, 1
        out.append(Request(t, kind, L, mode, nsw, nT, seed_base + i))
    return out
