"""Serving independent ground-state solves on one GPU.

A request is one NN_HUBBARD ground-state solve: a chain (makenetwork.py L
DMRG) or a T3NS tree (makenetwork.py L) of L sites at half filling, t = 1
and a given U, U(1) x U(1) symmetry, bond dimension D, a fixed number of
two-site sweeps (the reference runs one such calculation per process,
src/optimize_network.c:640-660; its globals become one Context per solve
here).  The server admits requests as they arrive while a solve slot is free
and the request's HBM footprint fits in what the running solves leave of
the budget, runs every admitted solve on its own host thread (its own
Context, HIP streams and device buffers) or, with a WorkerPool, in its own
worker process, so the work of concurrent solves overlaps on the device,
and records arrival, admission and completion times.  Requests that do not
fit wait; a later, smaller request may be admitted before them when it fits
(no slot is held idle).

Every solve is deterministic and independent of what runs beside it (no
shared device state; fixed-order reductions), so a result in the mix equals
the same request solved alone.
"""
import collections
import os
import threading
import time

import numpy as np

from .heff import Context
from .sweep import DeviceSolver, Sweeper

NETW_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                        "golden", "networks")

Request = collections.namedtuple("Request", ["rid", "geometry", "L", "U", "D", "sweeps", "arrival"])


def hubbard_stream(n, rate, seed=3, L_range=(8, 32), D_choices=(64, 128, 256, 512), sweeps=4):
    """The served-solve stream: n requests, L ~ U{L_range}, chain or tree
    with probability 1/2, U ~ U[1, 8], D uniform over D_choices, Poisson
    arrivals at `rate` per second (exponential gaps), all from one PCG64
    generator seeded `seed`."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        L = int(rng.integers(L_range[0], L_range[1] + 1))
        geo = "chain" if rng.random() < 0.5 else "tree"
        U = float(rng.uniform(1.0, 8.0))
        D = int(D_choices[int(rng.integers(len(D_choices)))])
        out.append(Request(i, geo, L, U, D, sweeps, t))
    return out


def footprint_bytes(req, max_vecs=30):
    """Upper estimate of a solve's device memory: the two-site object of a
    step holds at most D^2 x 16 states (two physical sites of dimension 4
    between bonds of at most D states); Davidson keeps 2 max_vecs + 3 such
    vectors, the Heff its operator arenas (at most 3 legs x 16 operator
    types of D x D) and work lists of the same order."""
    n = 16.0 * req.D * req.D
    return int(8 * (2 * max_vecs + 3) * n + 8 * 3 * 16 * req.D * req.D * 4 + 64 * n)


def solve(req, precision="f64"):
    """One request, start to finish, on the calling thread: returns the
    energy (the reference's minimum over sweeps) and the sweep energies."""
    name = "%d_%s" % (req.L, "DMRG" if req.geometry == "chain" else "T3NS")
    with open(os.path.join(NETW_DIR, name + ".netw")) as f:
        netw = f.read()
    ctx = Context(netw, "NN_HUBBARD", [1.0, req.U], ["U1", "U1"],
                  [(req.L + 1) // 2, req.L // 2], req.D)
    try:
        dm = Sweeper(ctx, req.D, seed=req.rid, solver=DeviceSolver(precision), tol=1e-8)
        e = dm.run(req.sweeps)
        return e, list(dm.energies)
    finally:
        ctx.close()


def _worker_init():
    # one HIP context per worker process, made before the stream starts
    from ._lib import lib
    lib().t3ns_device_count()


class WorkerPool:
    """Solves in worker processes (a solve's host work -- the sweep loop,
    planning, the small eigenproblems -- then runs on its own core instead of
    sharing one interpreter with the other solves; the device work of all
    workers overlaps on the one GPU).  Use as a context manager and pass as
    Server's solve_fn."""

    def __init__(self, workers):
        import concurrent.futures as cf
        import multiprocessing as mp
        self.pool = cf.ProcessPoolExecutor(max_workers=int(workers),
                                           mp_context=mp.get_context("spawn"),
                                           initializer=_worker_init)
        # start every worker now (their start-up is not part of a request)
        list(self.pool.map(_ping, range(int(workers))))

    def __call__(self, req, precision):
        return self.pool.submit(solve, req, precision).result()

    def close(self):
        self.pool.shutdown(wait=True)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _ping(i):
    return i


class Server:
    """Continuous admission of requests onto one GPU (see module doc)."""

    def __init__(self, slots=8, hbm_budget=200 << 30, precision="f64", solve_fn=None):
        """slots: solves running at once; hbm_budget: bytes of device memory
        the running solves may hold together (footprint_bytes each);
        solve_fn(request, precision) -> (energy, sweep energies), `solve`
        by default."""
        self.slots, self.budget, self.precision = int(slots), int(hbm_budget), precision
        self.solve_fn = solve_fn or solve

    def run(self, requests, clock=time.perf_counter):
        """Serves `requests` (arrival times in seconds from the start);
        returns one record per request: energy, sweep energies, arrival,
        admission and completion times, footprint."""
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            if footprint_bytes(r) > self.budget:
                raise ValueError("request %d needs more HBM than the whole budget" % r.rid)
        results = {}
        lock = threading.Condition()
        state = {"running": 0, "used": 0, "error": None}
        t0 = clock()
        waiting = collections.deque()
        threads = []

        def work(req, t_admit, fp):
            try:
                e, es = self.solve_fn(req, self.precision)
                rec = {"energy": e, "sweep_energies": es}
            except Exception as exc:  # reported after the run
                rec = {"error": repr(exc)}
                with lock:
                    state["error"] = state["error"] or exc
            t_done = clock() - t0
            with lock:
                rec.update({"rid": req.rid, "arrival": req.arrival, "admitted": t_admit,
                            "concurrent": state["running"], "used_at_end": state["used"],
                            "done": t_done, "latency": t_done - req.arrival,
                            "footprint": fp, "L": req.L, "D": req.D, "geometry": req.geometry})
                results[req.rid] = rec
                state["running"] -= 1
                state["used"] -= fp
                lock.notify_all()

        def admit_ready(now):
            """Starts every waiting request that fits, in arrival order."""
            i = 0
            while i < len(waiting) and state["running"] < self.slots:
                req = waiting[i]
                fp = footprint_bytes(req)
                if state["used"] + fp <= self.budget:
                    del waiting[i]
                    state["running"] += 1
                    state["used"] += fp
                    th = threading.Thread(target=work, args=(req, now, fp), daemon=True)
                    threads.append(th)
                    th.start()
                else:
                    i += 1

        k = 0
        with lock:
            while k < len(pending) or waiting or state["running"]:
                now = clock() - t0
                while k < len(pending) and pending[k].arrival <= now:
                    waiting.append(pending[k])
                    k += 1
                admit_ready(now)
                if k < len(pending):
                    timeout = max(0.0, pending[k].arrival - (clock() - t0))
                    lock.wait(timeout=timeout)
                elif waiting or state["running"]:
                    lock.wait()
        for th in threads:
            th.join()
        if state["error"] is not None:
            raise RuntimeError("a served solve failed") from state["error"]
        return [results[r.rid] for r in sorted(requests, key=lambda r: r.rid)]


def latency_summary(records, makespan=None):
    """solves/s, p50/p99 latency (seconds) of served records."""
    lat = np.array([r["latency"] for r in records])
    end = max(r["done"] for r in records)
    start = min(r["arrival"] for r in records)
    span = makespan if makespan is not None else end - start
    return {"solves": len(records), "solves_per_s": len(records) / span if span > 0 else None,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "makespan_s": span}
