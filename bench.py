"""Benchmark: 3D +-J Edwards-Anderson spin glass, L=128, parallel tempering over 32
temperatures geometric in [0.5, 2.0], 10 sweeps between replica-exchange attempts,
64 disorder samples per uint64 site word, on one MI355X.

    python bench.py [--gpus 1] [--steps 500] [--warmup 5]

A step is one parallel-tempering step of the whole ladder: 10 sweeps of all 32
temperatures (counter RNG), the energy measurement and one exchange attempt
decided on the host.  The default 500 steps are the full 5,000-sweep workload.
Prints one JSON line: replica spin-flip attempts/ns, the roofline of the dominant
kernel (HIP events on its stream) and a CPU baseline (the reference's own CPU
code, compiled from its sources, timed on this host).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

KIND, L, NT, SWEEPS = 2, 128, 32, 10
T_LADDER = [0.5 * 4.0 ** (k / (NT - 1)) for k in range(NT)]
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sweep3d_fused_L128_pt32_traffic.json")


class HipEvents:
    """hipEvent timing on a given stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def roofline_kernel(fused: bool) -> str:
    return "k_sweep3d_fused" if fused else "k_sweep3d_half"


def cpu_baseline():
    """The reference's parallel_tempering<cpu_sys> over cpu_3d (compiled from the
    reference sources into oracle/_ref), OpenMP over the temperatures as in the
    reference, on this host's cores: one step of 2 sweeps of the same ladder; and
    one sweep of one system on one core (`one_core`: the reference's sweep is
    serial)."""
    import oracle

    if oracle.ref_available():
        R = oracle.ref()
        cores = len(os.sched_getaffinity(0))
        sweeps = 2
        secs = R.ref_time_pt(KIND, L, NT, np.array(T_LADDER), 1234, sweeps, 1, cores)
        # the reference's sweep is serial: one system on one core as well
        one = R.ref_time_sweeps(KIND, L, 1.0, 1234, 1)
        return dict(value=NT * sweeps * L**3 * 64 / secs / 1e9, unit="attempts/ns", cores=cores, kind="reference",
                    sample=f"1 PT step x {sweeps} sweeps of the 32-temperature L=128 ladder (reference cpu_3d, "
                           "OpenMP)",
                    one_core=dict(value=L**3 * 64 / one / 1e9, unit="attempts/ns", cores=1,
                                  sample="1 sweep of one L=128 system at T=1 (reference cpu_3d)"))
    # the oracle restatement, one system on one core
    jx, jy, jz, s, g = oracle.init_system(KIND, L, 1234)
    tab = oracle.boltz(KIND, 1.0)
    w = g.fill(s.size)
    t0 = time.perf_counter()
    oracle.sweep(KIND, L, s, jx, jy, jz, tab, w)
    secs = time.perf_counter() - t0
    return dict(value=L**3 * 64 / secs / 1e9, unit="attempts/ns", cores=1, kind="port",
                sample="1 sweep of one L=128 system (oracle restatement, 1 core)")


def batched_vs_plain(sa, n=512):
    """The batched form against the plain path on the same instances.

    Workload: n independent realisations of uneven size (60% 2D periodic with L in
    {32,48,64,96,128}, 40% 3D with L in {16,24,32,48}), each its own linear anneal
    3.0 -> 0.1 over 500, 1000 or 2000 sweeps, counter RNG.  Batched: one Batch
    (one launch per half-sweep for all unfinished instances, hipGraph chunks, block
    size and chunk length tuned online inside the timed region).  Plain: a ladder
    per instance driven by sa_ladder_anneal (temperature upload + sweep launch per
    sweep, each instance on its own stream; 3D instances on the fused sweep, one
    launch per sweep, which is faster here than the two z-parallel launches a lone
    small ladder takes by default).  Set-up (seeding, couplings) is outside
    both timed regions.  Every instance's final spins are compared bit for bit."""
    from simulated_annealing_amd import workloads

    W = workloads.uneven_instances(n)
    attempts = sum(w.sites * 64 * w.nsweeps for w in W)
    b = sa.Batch()
    ids = [b.add(w.kind, w.L, w.seed, w.nsweeps, schedule=workloads.anneal_schedule(w.nsweeps)) for w in W]
    t0 = time.perf_counter()
    assert b.run() == 0
    dt_b = time.perf_counter() - t0
    lads = []
    for w in W:
        lad = sa.Ladder(w.kind, w.L, 1)
        lad.set_rng(sa.RNG_COUNTER)
        lad.seed_like_main(w.seed)
        if w.kind == sa.KIND_3D:  # one launch per sweep: the plain path's fastest form here
            lad.tune(sa.TUNE_SWEEP3D_KERNEL, 2)
        lads.append(lad)
    for lad in lads:
        lad.sync()
    t0 = time.perf_counter()
    for lad, w in zip(lads, W):
        lad.anneal(workloads.anneal_schedule(w.nsweeps))
    for lad in lads:
        lad.sync()
    dt_p = time.perf_counter() - t0
    same = sum(int((b.spins(i) == lad.spins(0)).all()) for i, lad in zip(ids, lads))
    # compulsory bytes of the batched run: per site-sweep 16 B of spins (read + write)
    # and the site's 2 (2D) or 3 (3D) coupling words, one system per instance
    alg = sum(w.sites * w.nsweeps * (16 + 8 * (3 if w.kind == sa.KIND_3D else 2)) for w in W)
    out = dict(workload=f"{n} uneven instances: 60% 2D L in {{32..128}}, 40% 3D L in {{16..48}}, "
                        "linear anneal 3.0->0.1 over 500/1000/2000 sweeps",
               batched_roofline_frac=round(alg / dt_b / (HBM_PEAK_GBS * 1e9), 4), alg_bytes=alg,
               attempts=attempts, batched_attempts_per_ns=round(attempts / dt_b / 1e9, 1),
               plain_attempts_per_ns=round(attempts / dt_p / 1e9, 1), speedup=round(dt_p / dt_b, 2),
               batched_s=round(dt_b, 4), plain_s=round(dt_p, 4), bitexact_instances=f"{same}/{n}",
               block=b.tuning(sa.TUNE_BATCH_BLOCK), chunk=b.tuning(sa.TUNE_BATCH_CHUNK))
    for lad in lads:
        lad.close()
    b.close()
    return out


def single_lattice(sa, L1=64, sweeps=2000):
    """One 3D L=64 lattice at T=1.0 (64 samples per word, counter RNG, seeded like
    main): sweeps of a single system, which offers L*L/2 columns per half-sweep.
    Roofline basis: 40 B per word-sweep (16 B spins read + written, 24 B couplings)."""
    lad = sa.Ladder(sa.KIND_3D, L1, 1)
    lad.set_rng(sa.RNG_COUNTER)
    lad.seed_like_main(1234)
    lad.set_temperature(0, 1.0)
    lad.sweep(50)
    lad.sync()
    t0 = time.perf_counter()
    lad.sweep(sweeps)
    lad.sync()
    dt = time.perf_counter() - t0
    N1 = L1**3
    tune = {k: lad.tuning(v) for k, v in (("kernel", sa.TUNE_SWEEP3D_KERNEL), ("rows_per_tile", sa.TUNE_ROWS_PER_TILE))}
    lad.close()
    us = dt / sweeps * 1e6
    return dict(workload=f"3D +-J EA L={L1}, one system at T=1.0, {sweeps} sweeps, counter RNG",
                attempts_per_ns=round(N1 * 64 * sweeps / dt / 1e9, 1), us_per_sweep=round(us, 2),
                roofline_frac=round(N1 * 40 / (us * 1e-6) / (HBM_PEAK_GBS * 1e9), 4), tuning=tune)


def serving(sa, n=1000, rate=40.0, budget=0, max_active=0):
    """The serving form: a Poisson stream of anneal / parallel-tempering requests.

    Requests (simulated_annealing_amd.workloads.serving_requests): 2D L log-uniform
    in [16, 256] or 3D L log-uniform in [8, 64] (odd L included), sweeps log-uniform in
    [100, 10000], half of them PT with 8..32 temperatures geometric in [0.5, 2.0]
    exchanging every 10 sweeps, the others a linear anneal 3.0 -> 0.1; per-request
    seeds.  All requests are submitted up front with their arrival times; the
    server admits each when it has arrived and memory / a slot is free.  Latency =
    result ready (spins digested, energies read back) - arrival."""
    from simulated_annealing_amd import workloads

    W = workloads.serving_requests(n, rate=rate)
    srv = sa.Server(hbm_budget=budget, max_active=max_active)
    base = srv.clock() + 0.25
    for w in W:
        srv.submit(sa.make_request(w.kind, w.L, w.mode, w.nsweeps, w.seed, nT=w.nT, every=10,
                                   arrival=base + w.arrival))
    # run in segments so a long stream reports progress (the loop keeps its clock)
    seg = max(1, n // 20)
    for m in list(range(seg, n, seg)) + [n]:
        srv.run(m)
        print(f"[serve] {m}/{n} finished at {srv.clock():.1f} s", file=sys.stderr, flush=True)
    res = srv.results()
    st = srv.stats()
    srv.close()
    ok = [x for x in res if x.status == 0]
    lat = np.array([x.finished - x.arrival for x in ok])
    wait = np.array([x.admitted - x.arrival for x in ok])
    span = max(x.finished for x in ok) - min(x.arrival for x in ok)
    attempts = sum(x.attempts for x in ok)
    reduced = "" if n >= 20000 else f" (the first {n} of the 20,000-request stream, so the default run ends in minutes)"
    return dict(workload=f"{n} Poisson requests at {rate:g}/s: 2D L~logU[16,256] / 3D L~logU[8,64], "
                         f"anneal or PT (8..32 T), sweeps~logU[100,10000]{reduced}",
                requests=n, completed=len(ok), rate_per_s=rate,
                p50_latency_s=round(float(np.percentile(lat, 50)), 4),
                p99_latency_s=round(float(np.percentile(lat, 99)), 4),
                mean_latency_s=round(float(lat.mean()), 4),
                p99_queue_wait_s=round(float(np.percentile(wait, 99)), 4),
                attempts_per_ns=round(attempts / span / 1e9, 1), span_s=round(span, 3),
                gpu_busy_frac=round(st["busy_s"] / span, 3), peak_active=st["peak_active"],
                peak_bytes=st["peak_bytes"], hbm_budget=st["hbm_budget"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-single", action="store_true", help="skip the single-lattice L=64 measurement")
    ap.add_argument("--serve-requests", type=int, default=1000, help="requests in the serving measurement")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-only", action="store_true", help="print only the serving measurement")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only the masked-swap kernel on two 128 MiB systems (known traffic: 256 MiB read, "
                         "256 MiB written per launch) to scale the FETCH_SIZE/WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        import simulated_annealing_amd as sa

        lad = sa.Ladder(KIND, 256, 2)
        for _ in range(5):
            lad.swap(0, 1, 0xFFFFFFFFFFFFFFFF)
        lad.sync()
        print(json.dumps({"calibration": "k_swap", "bytes_read_per_launch": 2 * 256**3 * 8,
                          "bytes_written_per_launch": 2 * 256**3 * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(independent ladders and batch instances shard across GPUs without communication: one process "
              "per GPU, each with its own share)", file=sys.stderr)
        sys.exit(2)

    import simulated_annealing_amd as sa

    if args.serve_only:
        print(json.dumps({"serving": serving(sa, args.serve_requests, args.serve_rate)}), flush=True)
        return
    pt = sa.PT(KIND, L, T_LADDER, 0.0, 1234, SWEEPS, sa.RNG_COUNTER)
    pt.step(args.warmup)
    pt.sync()
    t0 = time.perf_counter()
    pt.step(args.steps)
    pt.sync()
    dt = time.perf_counter() - t0
    N = L**3
    attempts = args.steps * SWEEPS * NT * N * 64
    value = attempts / dt / 1e9

    # dominant kernel: the 3D sweep over the whole ladder (one fused launch per
    # sweep, or two half-sweep launches).  Time back-to-back sweeps with hipEvents
    # on the ladder's stream.
    lad = pt.ladder_handle()
    stream = sa.lib().sa_ladder_stream(lad)
    v = C.c_int()
    sa.check(sa.lib().sa_ladder_get_tuning(lad, sa.TUNE_SWEEP3D_KERNEL, C.byref(v)), "tuning")
    fused = v.value == 2
    launches = 1 if fused else 2
    ev = HipEvents()
    a, b = ev.event(), ev.event()
    nsw = 40
    ev.record(a, stream)
    sa.check(sa.lib().sa_ladder_sweep(lad, nsw), "sweep")
    ev.record(b, stream)
    ms = ev.elapsed_ms(a, b)
    launch_us = ms * 1e3 / (launches * nsw)
    # algorithmic bytes per sweep: every spin word of the 32 systems read and written
    # once (16 B per word-sweep) plus each coupling word once for the whole ladder
    # (24 B per site, shared by the 32 systems); divided over the sweep's launches.
    alg_bytes = (NT * N * 16 + N * 24) / launches
    achieved = alg_bytes / (launch_us * 1e-6) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            prof = json.load(f)
        if prof.get("kernel") == roofline_kernel(fused):
            traffic = prof.get("hbm_bytes_per_launch")
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    kernel=roofline_kernel(fused), launch_us=round(launch_us, 2), alg_bytes_per_launch=alg_bytes)
    pt.close()
    single = None if args.no_single else single_lattice(sa)
    batched = None if args.no_batch else batched_vs_plain(sa)
    served = None if args.no_serve else serving(sa, args.serve_requests, args.serve_rate)
    cpu = None if args.no_cpu else cpu_baseline()
    line = {
        "metric": "replica spin-flip attempts/ns (1 GPU, 3D L=128); batched vs plain speedup",
        "value": round(value, 1),
        "unit": "attempts/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "3D +-J EA L=128 parallel tempering, 32 temperatures geometric 0.5..2.0, "
                               "10 sweeps per exchange, 64 samples per word",
                   "L": L, "temperatures": NT, "sweeps_per_step": SWEEPS, "rng": "counter (Philox4x32-10)",
                   "sweeps_timed": args.steps * SWEEPS},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "single_lattice": single,
        "batched_vs_plain": batched,
        "serving": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
