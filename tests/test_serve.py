"""Served solves (t3ns_amd.serve): admission by slots and HBM footprint on
the CPU with a stand-in solve, and on the GPU every served result equal to
the same request solved alone (the alone-vs-in-the-mix check), plus the
exact energy of a small chain."""
import threading
import time

import pytest

from t3ns_amd import serve
from t3ns_amd.serve import Request, Server, footprint_bytes, hubbard_stream, latency_summary


def test_stream_is_seeded_and_in_range():
    a = hubbard_stream(64, rate=2.0, seed=3)
    b = hubbard_stream(64, rate=2.0, seed=3)
    assert a == b
    assert all(8 <= r.L <= 32 and r.D in (64, 128, 256, 512) and 1 <= r.U <= 8 for r in a)
    assert {r.geometry for r in a} == {"chain", "tree"}
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))


def test_admission_respects_slots_and_budget():
    """A stand-in solve holds its slot for a while; the server never runs
    more than `slots` solves or more than the budget at once, a large
    request waits while smaller later ones that fit go ahead, and every
    request completes."""
    live = {"n": 0, "mem": 0, "max_n": 0, "max_mem": 0}
    lk = threading.Lock()
    order = []

    def fake(req, precision):
        fp = footprint_bytes(req)
        with lk:
            live["n"] += 1
            live["mem"] += fp
            live["max_n"] = max(live["max_n"], live["n"])
            live["max_mem"] = max(live["max_mem"], live["mem"])
            order.append(req.rid)
        time.sleep(0.05)
        with lk:
            live["n"] -= 1
            live["mem"] -= fp
        return -float(req.rid), [-float(req.rid)]

    reqs = [Request(0, "chain", 8, 4.0, 64, 1, 0.0),
            Request(1, "chain", 8, 4.0, 512, 1, 0.001),   # needs most of the budget
            Request(2, "chain", 8, 4.0, 64, 1, 0.002),
            Request(3, "tree", 8, 4.0, 64, 1, 0.003)]
    budget = footprint_bytes(reqs[1]) + footprint_bytes(reqs[0]) // 2
    srv = Server(slots=2, hbm_budget=budget, solve_fn=fake)
    recs = srv.run(reqs)
    assert [r["energy"] for r in recs] == [0.0, -1.0, -2.0, -3.0]
    assert live["max_n"] <= 2 and live["max_mem"] <= budget
    # request 1 cannot start beside request 0, so 2 goes first
    assert order.index(2) < order.index(1)
    s = latency_summary(recs)
    assert s["solves"] == 4 and s["p99_s"] >= s["p50_s"] > 0


def test_oversized_request_is_refused():
    srv = Server(slots=1, hbm_budget=1 << 20, solve_fn=lambda r, p: (0.0, [0.0]))
    with pytest.raises(ValueError):
        srv.run([Request(0, "chain", 8, 4.0, 64, 1, 0.0)])


@pytest.mark.gpu
def test_served_results_equal_alone(gpu):
    """Eight small requests (chains and trees, L 8-10, D 16-48, two sweeps)
    arriving 10 ms apart on three slots: each served energy and every sweep
    energy equals the same request solved alone to 1e-10 (the solves share
    nothing on the device)."""
    import numpy as np
    rng = np.random.default_rng(3)
    reqs = []
    for i in range(8):
        reqs.append(Request(i, "chain" if i % 2 == 0 else "tree", int(rng.integers(8, 11)),
                            float(rng.uniform(1, 8)), int(rng.choice([16, 32, 48])), 2, 0.01 * i))
    print("served requests", [(r.geometry, r.L, r.D) for r in reqs], flush=True)
    recs = Server(slots=3).run(reqs)
    assert max(r["concurrent"] for r in recs) >= 1
    for req, rec in zip(reqs, recs):
        e, es = serve.solve(req)
        assert abs(rec["energy"] - e) <= 1e-10 * abs(e), (req, rec["energy"], e)
        assert np.allclose(rec["sweep_energies"], es, rtol=1e-10, atol=0)


@pytest.mark.gpu
def test_served_chain_reaches_exact_energy(gpu):
    """A served 8-site chain at D = 256 (nothing truncated) gives the
    exact-diagonalisation energy."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from env_util import hubbard_ground_energy
    req = Request(0, "chain", 8, 4.0, 256, 3, 0.0)
    rec = Server(slots=1).run([req])[0]
    exact = hubbard_ground_energy(8, [(i, i + 1) for i in range(7)], 4, 4, 1.0, 4.0)
    assert abs(rec["energy"] - exact) < 1e-9, (rec["energy"], exact)


@pytest.mark.gpu
def test_served_in_worker_processes_equal_alone(gpu):
    """The same through worker processes (WorkerPool): results equal the
    solves run alone in this process."""
    reqs = [Request(i, ("chain", "tree")[i % 2], 8 + i, 3.0 + i, 32, 2, 0.005 * i) for i in range(4)]
    print("worker-pool requests", [(r.geometry, r.L) for r in reqs], flush=True)
    with serve.WorkerPool(2) as pool:
        recs = Server(slots=2, solve_fn=pool).run(reqs)
    for req, rec in zip(reqs, recs):
        e, _ = serve.solve(req)
        assert abs(rec["energy"] - e) <= 1e-10 * abs(e), (req, rec["energy"], e)
