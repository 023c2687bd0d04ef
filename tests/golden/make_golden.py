"""Writes tests/golden/ref_vectors.json from the reference's own compiled CPU classes
and drivers (oracle/_ref/libref.so, built by oracle/build_ref.sh from the reference
sources).  Run in the container that has the reference sources:

    python tests/golden/make_golden.py                 # ref_vectors.json
    python tests/golden/make_golden.py --checkpoints   # ckpt_vectors.json + ckpt_*.tempsys
    python tests/golden/make_golden.py --pt-files      # pt_files.json
    python tests/golden/make_golden.py --ground        # pt_ground.json
    python tests/golden/make_golden.py --configs       # config_vectors.json
    python tests/golden/make_golden.py --jfiles        # jfile_k*_L*.J (the reference's J files)

The fixtures are data only: inputs, seeds and the reference's outputs.  Spin arrays
are stored whole (hex words) for small lattices and as SHA-256 of their
little-endian bytes otherwise.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_vectors.json")
CKPT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ckpt_vectors.json")
PTF_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pt_files.json")
GROUND_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pt_ground.json")


def hexw(a) -> list[str]:
    return [f"{int(v):016x}" for v in np.asarray(a, dtype=np.uint64)]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def f32bits(a) -> list[str]:
    return [f"{int(v):08x}" for v in np.asarray(a, dtype=np.float32).view(np.uint32)]


def f64bits(a) -> list[str]:
    return [f"{int(v):016x}" for v in np.asarray(a, dtype=np.float64).view(np.uint64)]


def engines():
    R = O.ref()
    out = {}
    for seed in (5489, 1234, 42):
        a = np.zeros(10000, dtype=np.uint64)
        R.ref_mt64(seed, a, a.size)
        b = np.zeros(10000, dtype=np.uint32)
        R.ref_mt32(seed, b, b.size)
        u = np.zeros(16, dtype=np.float64)
        R.ref_uniform_real64(seed, u, u.size)
        out[str(seed)] = dict(mt64_first=hexw(a[:16]), mt64_10000=f"{int(a[-1]):016x}",
                              mt32_first=[int(x) for x in b[:16]], mt32_10000=int(b[-1]),
                              uniform_real64_first=f64bits(u))
    pts = {}
    for nT in (2, 3, 8, 32):
        p = np.zeros(200, dtype=np.int32)
        us = np.zeros(200, dtype=np.float64)
        R.ref_pt_draws(nT, p, us, p.size)
        pts[str(nT)] = dict(pairs=[int(x) for x in p], us=f64bits(us))
    return dict(engines=out, pt_draws=pts)


def tables():
    out = []
    for kind, T, h in [(2, 1.0, 0.0), (2, 0.5, 0.0), (2, 2.0, 0.0), (2, 0.7, 0.37), (2, 0.01, 0.0),
                       (1, 3.0, 0.0), (1, 0.1, 0.0), (1, 1.3, 0.25), (0, 1.3, 0.0), (0, 0.4, -0.2)]:
        r = O.RefSystem(kind, 4, T, h)
        out.append(dict(kind=kind, T=T, h=h, boltz=f64bits(r.state()["boltz"])))
    return out


def sweeps():
    cases = []
    for kind, L, T, h, seed, nsw in [
        (2, 8, 1.0, 0.0, 1234, 10), (2, 6, 0.8, 0.0, 7, 12), (2, 5, 1.5, 0.0, 99, 8), (2, 7, 1.1, 0.3, 5, 6),
        (2, 16, 1.0, 0.0, 1234, 4), (2, 4, 2.0, 0.0, 3, 20),
        (1, 16, 1.3, 0.0, 1234, 12), (1, 9, 0.9, 0.0, 11, 15), (1, 64, 2.0, 0.0, 1234, 3),
        (0, 16, 1.3, 0.0, 1234, 12), (0, 11, 0.6, 0.2, 17, 9), (0, 8, 1.0, 0.0, 2, 20),
    ]:
        r = O.RefSystem(kind, L, T, h)
        r.seed_like_main(seed)
        st0 = r.state()
        E0 = r.measure()
        recs = []
        for k in range(nsw):
            r.sweep(1)
            recs.append(f32bits(r.measure()))
        st = r.state()
        c = dict(kind=kind, L=L, T=T, h=h, seed=seed, sweeps=nsw, E0=f32bits(E0), E=recs,
                 s0_sha=sha(st0["s"]), s_sha=sha(st["s"]), jx_sha=sha(st0["Jx"][0::2]),
                 jy_sha=sha(st0["Jy"][0::2]))
        if kind == 2:
            c["jz_sha"] = sha(st0["Jz"][0::2])
        if st["s"].size <= 512:
            c["s"] = hexw(st["s"])
        cases.append(c)
    return cases


def temperature_schedule():
    """2D periodic L=16, per-sweep set_T on the linear schedule 3.0 -> 0.1, measure
    every 10 sweeps (the reference classes driven directly)."""
    kind, L, seed, n, stride = 1, 16, 1234, 200, 10
    Ts = [3.0 - 2.9 * k / (n - 1) for k in range(n)]
    r = O.RefSystem(kind, L, Ts[0])
    r.seed_like_main(seed)
    recs = []
    for k in range(n):
        r.set_T(Ts[k])
        r.sweep(1)
        if k % stride == 0:
            recs.append(f32bits(r.measure()))
    return dict(kind=kind, L=L, seed=seed, T=Ts, stride=stride, E=recs, s_sha=sha(r.state()["s"]))


def pt_runs():
    out = []
    for kind, L, nT, sweeps, nsteps, h in [(2, 6, 4, 3, 40, 0.0), (2, 8, 8, 10, 20, 0.0), (1, 16, 4, 5, 40, 0.0),
                                             (0, 12, 6, 2, 40, 0.0), (2, 6, 5, 2, 30, 0.3)]:
        T = [0.5 * 4 ** (k / (nT - 1)) for k in range(nT)][::-1]
        r = O.ref_pt_run(kind, L, np.array(T), h, 1234, sweeps, nsteps)
        out.append(dict(kind=kind, L=L, T=T, h=h, seed=1234, sweeps=sweeps, nsteps=nsteps,
                        pairs=[int(x) for x in r["pairs"]], us=f64bits(r["us"]), masks=hexw(r["masks"]),
                        accept=f64bits(r["accept"]), E=[f32bits(e) for e in r["E"]],
                        spins_sha=[sha(s) for s in r["spins"]]))
    return out


def anneal_geometric():
    out = []
    for kind, L, T0, alfa, sw, ns, stride in [(1, 16, 3.0, 0.9, 4, 10, 2), (2, 6, 2.5, 0.8, 3, 8, 1),
                                               (0, 12, 2.0, 0.95, 2, 6, 1)]:
        txt = O.ref_anneal_geometric(kind, L, T0, alfa, sw, ns, stride, 0.0, 1234)
        out.append(dict(kind=kind, L=L, T0=T0, alfa=alfa, sweeps=sw, nsteps=ns, stride=stride, seed=1234,
                        text=txt))
    return out


CKPT_FNAME = "sa_pt_%.dat"  # the file-name pattern the PT driver stores (sa_pt_create's default)


def checkpoints():
    """parallel_tempering::save records (.tempsys) written by the reference mid-run,
    stored as binary files beside this script, plus the run's continuation."""
    out = []
    here = os.path.dirname(os.path.abspath(__file__))
    for kind, L, nT, sweeps, n1, n2, h in [(2, 6, 4, 3, 5, 6, 0.0), (1, 8, 3, 4, 6, 5, 0.25), (0, 10, 4, 2, 4, 7, 0.0)]:
        T = [0.5 * 4 ** (k / (nT - 1)) for k in range(nT)]
        name = f"ckpt_k{kind}_L{L}.tempsys"
        r = O.ref_pt_checkpoint(kind, L, np.array(T), h, 4321, sweeps, n1, n2, CKPT_FNAME, os.path.join(here, name))
        out.append(dict(file=name, kind=kind, L=L, T=T, h=h, seed=4321, sweeps=sweeps, nsteps1=n1, nsteps2=n2,
                        fname=CKPT_FNAME, pairs=[int(x) for x in r["pairs"]], us=f64bits(r["us"]),
                        masks=hexw(r["masks"]), spins_sha=[sha(x) for x in r["spins"]]))
    return out


def pt_files():
    """Energy files, E_min file and print_swap text of the reference PT driver run
    with a stride (print(PRINT_E) every `stride` sweeps)."""
    import glob
    import tempfile

    out = []
    for kind, L, nT, sweeps, nsteps, stride, h in [(2, 6, 3, 6, 4, 2, 0.0), (1, 8, 4, 5, 3, 1, 0.25)]:
        T = [0.5 * 4 ** (k / (nT - 1)) for k in range(nT)]
        with tempfile.TemporaryDirectory() as d:
            swaps = O.ref_pt_files(kind, L, np.array(T), h, 555, sweeps, nsteps, stride, os.path.join(d, "pt_%.dat"))
            files = {os.path.basename(f): open(f).read() for f in sorted(glob.glob(os.path.join(d, "*")))}
        out.append(dict(kind=kind, L=L, T=T, h=h, seed=555, sweeps=sweeps, nsteps=nsteps, stride=stride,
                        swap_text=swaps, files=files))
    return out


def ground():
    """The ground-state loop pieces of main_cpu -r on one reference PT driver: rounds
    of steps, get_E_min, update_T, reset, then steps once more (replay stream)."""
    out = []
    for kind, L, T, h, seed, sweeps, nsteps, rounds in [(2, 8, [0.4, 1.2, 3.0], 0.0, 77, 5, 20, 2),
                                                         (1, 8, [0.3, 1.0, 2.5, 4.0], 0.2, 78, 4, 15, 2),
                                                         (0, 10, [0.5, 2.0, 5.0], 0.0, 79, 3, 12, 3)]:
        r = O.ref_pt_ground(kind, L, np.array(T), h, seed, sweeps, nsteps, rounds)
        out.append(dict(kind=kind, L=L, T=T, h=h, seed=seed, sweeps=sweeps, nsteps=nsteps, rounds=rounds,
                        nT=[int(x) for x in r["nT"]],
                        acc=[f64bits(r["acc"][q][:(T.__len__() if q == 0 else int(r["nT"][q - 1])) - 1])
                             for q in range(rounds)],
                        T_rounds=[f64bits(t) for t in r["T_rounds"]], E_min=[f32bits(e) for e in r["E_min"]],
                        spins_sha=[sha(x) for x in r["spins"]], E=[f32bits(e) for e in r["E"]]))
    return out


CONFIG_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config_vectors.json")


def config_runs():
    """The benchmark configurations at their stated sizes, from the reference's own
    classes and drivers (replay stream, seed 1234):
      * 2D periodic L=64, linear anneal T_k = 3.0 - 2.9 k/999 over 1,000 sweeps
        (set_T before every sweep), 64 energies after every sweep k with k % 10 == 0;
      * 3D L=64 at T=1.0, 10,000 sweeps: energies every 1,000 sweeps, final spins;
      * 3D L=128 parallel tempering over 32 temperatures geometric 0.5 .. 2.0,
        10 sweeps per exchange attempt, 3 steps: pair, uniform and mask per step,
        the energies of all 32 temperatures and every system's spins."""
    out = {}
    kind, L, seed, n, stride = 1, 64, 1234, 1000, 10
    Ts = [3.0 - 2.9 * k / (n - 1) for k in range(n)]
    r = O.RefSystem(kind, L, Ts[0])
    r.seed_like_main(seed)
    recs = []
    for k in range(n):
        r.set_T(Ts[k])
        r.sweep(1)
        if k % stride == 0:
            recs.append(f32bits(r.measure()))
    out["anneal2d_L64"] = dict(kind=kind, L=L, seed=seed, T=Ts, stride=stride, E=recs,
                               E_final=f32bits(r.measure()), s_sha=sha(r.state()["s"]))
    print("config 1 done", flush=True)
    kind, L, T, n, every = 2, 64, 1.0, 10000, 1000
    r = O.RefSystem(kind, L, T)
    r.seed_like_main(seed)
    recs, shas = [], []
    for k in range(n // every):
        r.sweep(every)
        recs.append(f32bits(r.measure()))
        shas.append(sha(r.state()["s"]))
        print(f"config 2: {(k + 1) * every} sweeps", flush=True)
    out["fixed3d_L64"] = dict(kind=kind, L=L, seed=seed, T=T, sweeps=n, every=every, E=recs, s_sha=shas)
    kind, L, nT, sweeps, nsteps = 2, 128, 32, 10, 3
    T = [0.5 * 4.0 ** (k / (nT - 1)) for k in range(nT)]
    p = O.ref_pt_run(kind, L, np.array(T), 0.0, seed, sweeps, nsteps)
    out["pt3d_L128"] = dict(kind=kind, L=L, T=T, h=0.0, seed=seed, sweeps=sweeps, nsteps=nsteps,
                            pairs=[int(x) for x in p["pairs"]], us=f64bits(p["us"]), masks=hexw(p["masks"]),
                            E=[f32bits(e) for e in p["E"]], spins_sha=[sha(s) for s in p["spins"]])
    print("config 3 done", flush=True)
    return out


JFILES = [(2, 6, 1234), (1, 8, 77), (0, 10, 5)]  # (kind, L, seed) of the J-file fixtures


def jfiles():
    """J files written by the reference's own save_J (src/sys_file.cpp:10-110) for a
    system seeded like main (set_seed, init_J): jfile_k<kind>_L<L>.J."""
    R = O.ref()
    for kind, L, seed in JFILES:
        p = R.ref_create(kind, L, 1.0, 0.0)
        R.ref_set_seed(p, seed)
        R.ref_init_J(p)
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"jfile_k{kind}_L{L}.J")
        R.ref_save_J(p, path.encode())
        R.ref_destroy(p)
        print("wrote", path, os.path.getsize(path), "bytes")


def main():
    if not O.ref_available():
        O.build_ref()
    if "--jfiles" in sys.argv:
        jfiles()
        return
    if "--configs" in sys.argv:  # only the benchmark-configuration fixtures
        with open(CONFIG_OUT, "w") as f:
            json.dump(dict(about="The benchmark configurations run by the reference (see make_golden.py).",
                           **config_runs()), f, indent=0, separators=(",", ":"))
        print("wrote", CONFIG_OUT)
        return
    if "--ground" in sys.argv:  # only the ground-state loop fixtures
        with open(GROUND_OUT, "w") as f:
            json.dump(dict(about="update_T / reset / get_E_min of the reference PT driver (see make_golden.py).",
                           runs=ground()), f, indent=0, separators=(",", ":"))
        print("wrote", GROUND_OUT)
        return
    if "--pt-files" in sys.argv:  # only the PT output-file fixtures
        with open(PTF_OUT, "w") as f:
            json.dump(dict(about="parallel_tempering output files of the reference (see make_golden.py).",
                           runs=pt_files()), f, indent=0, separators=(",", ":"))
        print("wrote", PTF_OUT)
        return
    if "--checkpoints" in sys.argv:  # only the checkpoint fixtures
        with open(CKPT_OUT, "w") as f:
            json.dump(dict(about="parallel_tempering::save records of the reference (see make_golden.py).",
                           checkpoints=checkpoints()), f, indent=0, separators=(",", ":"))
        print("wrote", CKPT_OUT)
        return
    doc = dict(
        about="Outputs of the reference's compiled CPU classes/drivers (see make_golden.py).",
        **engines(), tables=tables(), sweeps=sweeps(), schedule=temperature_schedule(), pt=pt_runs(),
        anneal=anneal_geometric())
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=0, separators=(",", ":"))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
