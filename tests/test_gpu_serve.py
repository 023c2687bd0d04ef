"""The serving loop (sa_serve_*): continuous batching of anneal / PT requests.

Every request's result (digest of all final spin words, final energies, PT
exchange log) must equal running that request alone: through the plain path
(a ladder seeded like main, sa_ladder_anneal / the PT driver with the counter
RNG) and through a server that holds only that request.  Admission must respect
the device-memory budget and the slot limit, in arrival order."""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def sa():
    import simulated_annealing_amd as sa

    sa.lib()
    return sa


def mixed_requests(sa, n=24, seed=3, rate=400.0):
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for q in range(n):
        t += rng.exponential(1.0 / rate)
        kind = int(rng.choice([0, 1, 2]))
        L = int(rng.choice([7, 8, 12, 13, 16] if kind == 2 else [8, 15, 16, 24, 31, 32]))  # odd L too
        h = float(rng.choice([0.0, 0.0, 0.25]))
        if rng.random() < 0.5:
            r = sa.make_request(kind, L, sa.REQ_ANNEAL, int(rng.integers(20, 200)), 7000 + q, h=h, arrival=t)
        else:
            every = int(rng.choice([5, 10]))
            r = sa.make_request(kind, L, sa.REQ_PT, every * int(rng.integers(2, 15)), 7000 + q,
                                nT=int(rng.integers(2, 9)), every=every, h=h, arrival=t)
        out.append(r)
    return out


def alone(sa, r):
    """(spins digest, E, M, log digest) of the request run alone through the plain path."""
    T = sa.request_temperatures(r)
    if r.mode == sa.REQ_ANNEAL:
        lad = sa.Ladder(r.kind, r.L, 1)
        lad.set_rng(sa.RNG_COUNTER)
        lad.seed_like_main(r.seed)
        lad.anneal(T, r.h)
        E, M = lad.measure()
        d = sa.digest(lad.spins(0), 0)
        lad.close()
        return d, E[0], M[0], 0
    pt = sa.PT(r.kind, r.L, T, r.h, r.seed, r.every, sa.RNG_COUNTER)
    pt.step(r.nsweeps // r.every)
    res = pt.result()
    N = res["spins"].shape[1]
    d = 0
    for j in range(r.nT):
        d ^= sa.digest(res["spins"][j], j * N)
    lad = pt.ladder_handle()
    E = np.zeros(r.nT * 64, dtype=np.int32)
    M = np.zeros(r.nT * 64, dtype=np.int32)
    sa.check(sa.lib().sa_ladder_measure(lad, E, M), "measure")
    ld = sa.log_digest(res["pairs"], res["us"], res["masks"])
    pt.close()
    return d, E[:64], M[:64], ld


def test_serve_results_equal_alone(sa):
    reqs = mixed_requests(sa)
    # tight memory: about three of the larger requests at a time, at most 6 slots
    budget = 3 * max(sa.request_bytes(r) for r in reqs)
    srv = sa.Server(hbm_budget=budget, max_active=6)
    say("serve", len(reqs), "requests, budget", budget)
    for r in reqs:
        srv.submit(r)
    srv.run()
    st = srv.stats()
    assert st["finished"] == len(reqs) and st["peak_bytes"] <= budget and st["peak_active"] <= 6
    res = srv.results()
    for r, x in zip(reqs, res):
        say("request", x.id, "mode", r.mode, "kind", r.kind, "L", r.L, "sweeps", r.nsweeps, "nT", r.nT, "h", r.h)
        assert x.status == 0
        assert x.arrival <= x.admitted + 1e-9 <= x.finished + 2e-9
        d, E, M, ld = alone(sa, r)
        assert x.spins_digest == d
        assert (np.array(x.E) == E).all() and (np.array(x.M) == M).all()
        if r.mode == sa.REQ_PT:
            assert x.log_digest == ld and x.steps == r.nsweeps // r.every
    # the same requests alone in a server of their own
    for q in (0, 5, 11):
        solo = sa.Server()
        r = reqs[q]
        r1 = sa.Request(r.kind, r.L, r.mode, r.nT, r.nsweeps, r.every, r.T0, r.T1, r.h, r.seed, 0.0)
        solo.submit(r1)
        solo.run()
        y = solo.results()[0]
        assert y.status == 0 and y.spins_digest == res[q].spins_digest and y.log_digest == res[q].log_digest
        solo.close()
    srv.close()


def test_serve_admission_order_and_refusals(sa):
    small = [sa.make_request(1, 16, sa.REQ_ANNEAL, 30, 100 + q, arrival=0.001 * q) for q in range(6)]
    big = sa.make_request(2, 32, sa.REQ_PT, 20, 99, nT=8, every=10, arrival=0.0025)
    tiny = sa.make_request(1, 2, sa.REQ_ANNEAL, 10, 98)
    fp = sa.request_bytes(small[0])
    assert sa.request_bytes(big) > 2 * fp + 1024
    budget = 2 * fp + 1024  # two small requests at a time; the PT request never fits
    srv = sa.Server(hbm_budget=budget, max_active=8)
    ids = [srv.submit(r) for r in small[:3]]
    ib = srv.submit(big)
    ids += [srv.submit(r) for r in small[3:]]
    with pytest.raises(sa.SAError):
        srv.submit(tiny)  # refused at submission: L < 3
    srv.run()
    res = {x.id: x for x in srv.results()}
    st = srv.stats()
    assert st["peak_bytes"] <= budget and st["peak_active"] <= 2
    assert res[ib].status == -4  # SA_ERR_NOMEM: over the budget on its own
    adm = [res[i].admitted for i in ids]
    assert adm == sorted(adm)  # arrival order
    for i in ids:
        assert res[i].status == 0
    srv.close()


def test_serve_online_submission_from_another_thread(sa):
    srv = sa.Server()
    reqs = [sa.make_request(2, 8, sa.REQ_ANNEAL, 50, 300 + q, arrival=-1.0) for q in range(5)]

    def client():
        for r in reqs:
            srv.submit(r)
            time.sleep(0.01)

    th = threading.Thread(target=client)
    th.start()
    srv.run(len(reqs))
    th.join()
    res = srv.results()
    assert len(res) == len(reqs) and all(x.status == 0 for x in res)
    for r, x in zip(reqs, res):
        d, E, M, _ = alone(sa, r)
        assert x.spins_digest == d
    srv.close()
