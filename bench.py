#!/usr/bin/env python3
"""Heff matvec benchmark on one MI355X.

Workload: synthetic quantum-chemistry Hamiltonian with 24 orbitals and 24
electrons (FCIDUMP from t3ns_amd.synthetic, seed 1, ERI rank 48), symmetry
U1 x SU2 (N = 24, 2S = 0), T3NS tree network with 24 physical sites,
bond dimension D = 512 (bookkeeper sector dims), three-site optimisation
step {physical, branching, physical} around branching site 31 (sites 18, 31,
17).  One step of the benchmark = one effective-Hamiltonian matvec on a
device-resident N(0,1) vector (seed 1); the operators hold N(0,1) entries on
the exact block structure of the renormalised operators of that step.

Prints one JSON line: matvec/s of the plain FP64 path (value) and of the
FP32 path, the roofline of the matvec kernel (HIP events on its stream), and
the CPU baseline (the FP64 restatement in oracle/, OpenMP over output blocks,
timed on a bounded sample of output blocks on this host).  The "solve" entry
is the ground-state solve of the same step on the Hamiltonian's own
operators (renormalised on the device from the vacuum inwards): plain FP64
Davidson against FP32 matvecs with FP64 refinement, energies, iterations,
seconds and |dE|.  The "sweep" entry is the two-site DMRG of the 10-site
NN_HUBBARD chain (t = 1, U = 4, half filling, D = 64, up to 4 sweeps,
E_CONV = 1e-10) with every step on the device: seconds and energies, FP64
and mixed; and the reference's LiF regression sweep (tests/test3.c.in,
singlet) with its reference energy.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# matrix-core peaks of the arithmetic each path runs (the BF16 path stores
# operators in BF16 and computes on FP32 matrix cores)
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 157.3}


def profiled_traffic(workload, sites, D):
    """HBM bytes per FP64 matvec from the committed PMC summary of this
    workload (FETCH_SIZE/WRITE_SIZE, copy-calibrated), or None."""
    prof = os.path.join(ROOT, "profiles", "heff_matvec_%s_t3ns_D%d.md" % (workload, D))
    if sites != WORKLOADS[workload]["sites"] or D != WORKLOADS[workload]["bond_dim"] or \
            not os.path.exists(prof):
        return None
    with open(prof) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_matvec:"):
                return float(line.split(":")[1])
    return None
NETWORK = os.path.join(ROOT, "tests", "golden", "networks", "24_T3NS.netw")
# The matvec workloads: synthetic FCIDUMP (orbitals = electrons, seed, ERI
# rank), U1 x SU2, the makenetwork.py tree, bond dimension and step sites.
WORKLOADS = {
    # three-site {physical, branching, physical} step of the QC-24 system
    "qc24": {"norb": 24, "seed": 1, "rank": 48, "bond_dim": 512, "sites": [18, 31, 17]},
    # two-site step 8 of the sweep of the QC-50 system at D = 2048: a
    # branching site whose two virtual bonds carry 2083 and 2066 reduced
    # states (29.5 M elements, 2.9e14 flops per matvec)
    "qc50": {"norb": 50, "seed": 4, "rank": 100, "bond_dim": 2048, "sites": [69, 42]},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="qc24",
                    help="qc24: the QC-24 T3NS D=512 three-site step (default, with every other "
                         "leg); qc50: the QC-50 T3NS D=2048 two-site step (matvec legs only)")
    ap.add_argument("--sites", default=None, help="step sites (default: the workload's)")
    ap.add_argument("--bond-dim", type=int, default=None, help="default: the workload's")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the baseline sample (0 disables it)")
    ap.add_argument("--precisions", default="f64,f32,bf16")
    ap.add_argument("--autotune-trials", type=int, default=6,
                    help="work-list settings the autotuner times before the timed matvecs of "
                         "each precision (0: the defaults, untuned)")
    ap.add_argument("--wpc", type=int, default=0, choices=(0, 1, 2, 4),
                    help="with --autotune-trials 0: workgroups per CU of the matvec kernels "
                         "(0: picked from the object's structure)")
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the renormalisation and site-tensor kernels")
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the ground-state solve on the Hamiltonian's own operators")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the DMRG sweep of the 10-site Hubbard chain")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the served-solve stream")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-matvec leg (15 small objects)")
    ap.add_argument("--qc24-sweep", type=int, default=0, metavar="D",
                    help="also run one two-site sweep of the QC-24 system at bond dimension D")
    ap.add_argument("--serve-requests", type=int, default=12,
                    help="requests of the served-solve stream (the first ones of the seed-3 stream)")
    ap.add_argument("--serve-rate", type=float, default=4.0,
                    help="Poisson arrival rate of the served-solve stream (requests/s)")
    ap.add_argument("--solve-tol", type=float, default=1e-5,
                    help="Davidson residual tolerance of the solve leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (scale reference for HBM counters)")
    return ap.parse_args()


def build_step(sites, D, workload="qc24"):
    import t3ns_amd as T
    from t3ns_amd import synthetic
    w = WORKLOADS[workload]
    n = w["norb"]
    fci = synthetic.qc_workload(n, n, seed=w["seed"], rank=w["rank"])
    with open(os.path.join(ROOT, "tests", "golden", "networks", "%d_T3NS.netw" % n)) as f:
        netw = f.read()
    ctx = T.Context(netw, "QC", [], ["U1", "SU2"], [n, 0], D, fcidump=fci)
    step = T.Step(ctx, sites, seed=1)
    return ctx, step


def cpu_baseline(step, x, target_seconds, flops_total, min_fraction=0.25, max_seconds=40.0):
    """FP64 oracle matvec over a strided sample of output blocks, its GEMMs
    on OpenBLAS (one BLAS thread per OpenMP thread, as the reference runs
    cblas_dgemm inside its OpenMP loop over output blocks).  The sample
    covers at least `min_fraction` of the matvec's flops when that fits in
    about max_seconds, else what `target_seconds` allows."""
    import oracle as O
    blas = O.use_blas()
    d, o, s = step.descriptors()
    t0 = time.perf_counter()
    oh = O.OracleHeff(d, o, s)
    print("# cpu baseline: oracle plan in %.1f s (BLAS %s)" % (time.perf_counter() - t0, blas),
          file=sys.stderr, flush=True)
    bf = oh.block_flops()
    out = np.zeros(oh.size)
    nb = oh.nrblocks
    # calibrate on a strided subset of about 2 s of work at a guessed
    # 5 GFLOP/s per thread (the OpenMP loop runs over the subset's blocks)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    want_cal = 2.0 * 5e9 * threads
    cal = nb
    while cal > 1 and bf[0:nb:cal].sum() < want_cal:
        cal = max(1, cal // 2)
    c0 = time.perf_counter()
    oh.matvec_blocks(x, out, 0, nb, cal)
    t_cal = time.perf_counter() - c0
    f_cal = bf[0:nb:cal].sum()
    rate = f_cal / max(t_cal, 1e-9)
    print("# cpu baseline: calibration %.1f s, %.3g flop/s" % (t_cal, rate), file=sys.stderr,
          flush=True)
    frac = min(1.0, rate * target_seconds / max(flops_total, 1.0))
    if rate * max_seconds >= min_fraction * flops_total:
        frac = max(frac, min_fraction)
    stride = max(1, int(1.0 / max(frac, 1e-9)))
    t0 = time.perf_counter()
    oh.matvec_blocks(x, out, 0, nb, stride)
    t = time.perf_counter() - t0
    f = bf[0:nb:stride].sum()
    oh.close()
    return {"value": (f / t) / flops_total, "unit": "matvec/s", "cores": threads,
            "kind": "port",
            "sample": "FP64 oracle restatement of src/Heff.c matvec (GEMMs on OpenBLAS %s), "
                      "OpenMP over output blocks, every %d-th output block (%d of %d blocks, "
                      "%.1f%% of the matvec flops) in %.1f s; matvec/s = sample flop rate / "
                      "full-matvec flops" % ("on" if blas else "unavailable: loop nest", stride,
                                             len(range(0, nb, stride)), nb,
                                             100.0 * f / flops_total, t)}


def cpu_sweep_baseline(sweeps=4, e_conv=1e-10):
    """The sweep leg's Hubbard chain (10 sites, D = 64, the same N(0, 1)
    seed-0 start) with every step on the oracle, the FP64 restatement of
    the reference's CPU path (OpenMP matvec, Lanczos eigensolver,
    renormalisation, split), on this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    import t3ns_amd as T
    from env_util import OracleSolver
    O.use_blas()
    from t3ns_amd.sweep import Sweeper
    with open(os.path.join(ROOT, "tests", "golden", "networks", "10_DMRG.netw")) as f:
        netw = f.read()
    ctx = T.Context(netw, "NN_HUBBARD", [1.0, 4.0], ["U1", "U1"], [5, 5], 64)
    t0 = time.perf_counter()
    dm = Sweeper(ctx, 64, seed=0, solver=OracleSolver(), tol=1e-5, start="normal")
    e = dm.run(sweeps, e_conv=e_conv)
    t = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"workload": "nn_hubbard_10_chain_D64", "energy": e, "sweeps": len(dm.energies),
            "seconds": t, "seconds_per_sweep": t / len(dm.energies), "cores": threads,
            "kind": "port"}


HBM_GBS = 8000.0
RIDGE = PEAK_TFLOPS["f64"] * 1e12 / (HBM_GBS * 1e9)  # FP64 flop per byte


def _row(flops, nbytes, ms, extra=None):
    """Roofline line of one kernel group: the bound is the FP64 matrix peak
    when its flop/byte ratio is above the ridge point, HBM otherwise."""
    s = max(ms, 1e-9) * 1e-3
    tf, gbs = flops / s / 1e12, nbytes / s / 1e9
    intensity = flops / max(nbytes, 1.0)
    if intensity >= RIDGE:
        r = {"bound": "mfma", "achieved": tf, "peak": PEAK_TFLOPS["f64"], "unit": "TFLOP/s",
             "frac": tf / PEAK_TFLOPS["f64"]}
    else:
        r = {"bound": "hbm", "achieved": gbs, "peak": HBM_GBS, "unit": "GB/s", "frac": gbs / HBM_GBS}
    r.update({"kernel_ms": ms, "flops": flops, "bytes": nbytes})
    if extra:
        r.update(extra)
    return r


def row_problems(step, bond_dim, sites):
    """(name, device call -> (flops, bytes, kernel ms), oracle call) for the
    renormalisation and site-tensor rows on one step object."""
    import oracle as O
    import t3ns_amd as T
    from t3ns_amd import synthetic
    out = []
    d, o, s = step.descriptors()

    def dev_rops(fn, *a):
        return lambda: fn(*a).work()

    for leg in range(3):
        if o[leg].nrhss <= 0 or not o[leg].P_operator:
            continue
        ad, old = step.append_problem(leg, seed=3)
        out.append(("append_phys_leg%d" % leg, dev_rops(T.append_phys, ad, old),
                    lambda ad=ad, old=old: O.append_phys(ad, old)))
        pd, pops = step.pupdate_problem(leg, seed=5, maxdim=0)
        out.append(("update_physical_leg%d" % leg, dev_rops(T.update_physical, pd, pops),
                    lambda pd=pd, pops=pops: O.update_physical(pd, pops)))
    md = step.makesite_problem(seed=7)
    out.append(("makesite", lambda: T.make_site_tensor(md)[1], lambda: O.make_site_tensor(md)))
    for sv in range(s.nrsites):
        sd = step.split_problem(sv, maxD=bond_dim)
        try:
            O.split_site(step.split_problem(sv, maxD=1))
        except RuntimeError:
            continue  # the branching site is not cut off by one bond
        out.append(("split_site%d" % sv, lambda sd=sd: T.split_site(sd)[1], lambda sd=sd: O.split_site(sd)))
    fci = synthetic.qc_workload(24, 24, seed=1, rank=48)
    with open(NETWORK) as fh:
        netw = fh.read()
    ctx = T.Context(netw, "QC", [], ["U1", "SU2"], [24, 0], bond_dim, fcidump=fci)
    branch = [s.sites[i] for i in range(s.nrsites) if not s.site_is_phys[i]]
    for uc in range(3 if branch else 0):
        bd, bo = ctx.bupdate_problem(branch[0], uc, seed=3)
        out.append(("update_branching_case%d" % uc, dev_rops(T.update_branching, bd, bo),
                    lambda bd=bd, bo=bo: O.update_branching(bd, bo)))
    qd = ctx.qr_problem(sites[0], sites[1], seed=3)
    out.append(("qr_step", lambda: T.qr_step(qd)[1], lambda: O.qr_step(qd)))
    return out, ctx


def measure_rows(step, sites, bond_dim, cpu_dim=128, same_size_flops=1e10):
    """The renormalisation and site-tensor rows on the benchmark object:
    every device call runs twice and the second is reported (kernel time
    from device events, algorithmic flops and bytes).  The CPU baseline of a
    row is the FP64 oracle restatement (GEMMs on OpenBLAS, the host's
    OpenMP threads where the restatement has them) on the same row of the
    same step: at the same size when the row has at most same_size_flops
    flops, else at bond dimension `cpu_dim` (the full size would take
    minutes); its rate is flops per second of the instance it ran."""
    import oracle as O
    O.use_blas()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rows = {}
    probs, ctx = row_problems(step, bond_dim, sites)
    for name, dev, ora in probs:
        dev()
        f, b, ms = dev()
        rows[name] = _row(f, b, ms)
        if f <= same_size_flops:
            t0 = time.perf_counter()
            ora()
            t = time.perf_counter() - t0
            rows[name]["cpu_baseline"] = {"value": f / max(t, 1e-9) / 1e9, "unit": "GFLOP/s",
                                          "seconds": t, "flops": f, "cores": threads,
                                          "kind": "port",
                                          "sample": "same row of the same step, same size"}
    ctx.close()
    if all("cpu_baseline" in r for r in rows.values()):
        return rows
    _, small = build_step(sites, cpu_dim)
    sprobs, sctx = row_problems(small, cpu_dim, sites)
    for name, dev, ora in sprobs:
        if name not in rows or "cpu_baseline" in rows[name]:
            continue
        f, _, _ = dev()
        t0 = time.perf_counter()
        ora()
        t = time.perf_counter() - t0
        rows[name]["cpu_baseline"] = {"value": f / max(t, 1e-9) / 1e9, "unit": "GFLOP/s",
                                      "seconds": t, "flops": f, "cores": threads, "kind": "port",
                                      "sample": "same row of the same step at D=%d (the row at "
                                                "D=%d has %.3g flops)" % (cpu_dim, bond_dim,
                                                                          rows[name]["flops"])}
    sctx.close()
    return rows


def _time_matvec(h, steps=3):
    """ms per device-resident matvec of a Heff handle (HIP events on the
    launch stream, one warm-up call first)."""
    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    dx = torch.from_numpy(np.random.default_rng(7).standard_normal(h.size)).to(dev)
    dy = torch.empty_like(dx)
    torch.cuda.synchronize()
    h.matvec_device(dx.data_ptr(), dy.data_ptr(), stream.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        h.matvec_device(dx.data_ptr(), dy.data_ptr(), stream.cuda_stream)
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_time(e1) / steps


def measure_batch(reps=20):
    """Cross-solve batching: the matvecs of the 15 two-site steps of a
    16-site NN_HUBBARD chain (t = 1, U = 4, U1 x U1 (8, 8), D = 128, seeded
    operators), as 15 small concurrent solves would issue them: one after
    another on one stream, each on its own stream, and in one batched
    launch per kernel.  ms per round of 15 matvecs (HIP events), and
    whether the batched results equal the separate ones bitwise."""
    import torch
    import t3ns_amd as T
    with open(os.path.join(ROOT, "tests", "golden", "networks", "16_DMRG.netw")) as f:
        netw = f.read()
    ctx = T.Context(netw, "NN_HUBBARD", [1.0, 4.0], ["U1", "U1"], [8, 8], 128)
    steps = [T.Step(ctx, [i, i + 1], seed=2) for i in range(15)]
    hs = [T.Heff.from_step(st, T.FP64) for st in steps]
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    xs = [torch.from_numpy(rng.standard_normal(h.size)).to(dev) for h in hs]
    ys = [torch.empty_like(x) for x in xs]
    streams = [torch.cuda.Stream(device=dev) for _ in hs]
    main = torch.cuda.Stream(device=dev)
    b = T.HeffBatch()

    def seq():
        for h, x, y in zip(hs, xs, ys):
            h.matvec_device(x.data_ptr(), y.data_ptr(), main.cuda_stream)

    def multi():
        ev = torch.cuda.Event()
        ev.record(main)
        for h, x, y, st in zip(hs, xs, ys, streams):
            st.wait_event(ev)
            h.matvec_device(x.data_ptr(), y.data_ptr(), st.cuda_stream)
        for st in streams:
            main.wait_stream(st)

    def batched():
        b.matvec_device(hs, [x.data_ptr() for x in xs], [y.data_ptr() for y in ys], main.cuda_stream)

    out = {"workload": "nn_hubbard_16_chain_D128_15_steps", "objects": len(hs),
           "sizes": [h.size for h in hs]}
    ref = None
    for name, fn in (("one_stream", seq), ("own_streams", multi), ("batched", batched)):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(main)
        for _ in range(reps):
            fn()
        e1.record(main)
        torch.cuda.synchronize()
        out[name + "_ms"] = e0.elapsed_time(e1) / reps
        res = [y.cpu().numpy().copy() for y in ys]
        if ref is None:
            ref = res
        else:
            out[name + "_equal"] = all(np.array_equal(a, r) for a, r in zip(res, ref))
    out["speedup_vs_one_stream"] = out["one_stream_ms"] / out["batched_ms"]
    out["speedup_vs_own_streams"] = out["own_streams_ms"] / out["batched_ms"]
    b.close()
    print("# batch: %s" % out, file=sys.stderr, flush=True)
    return out


def measure_solve(ctx, step, tol):
    """Lowest eigenpair of the step's effective Hamiltonian built from the
    Hamiltonian's own operators: every bond's operators renormalised on the
    device from the vacuum inwards (t3ns_amd.environment, init_rops of
    src/optimize_network.c:428-468) with seeded site tensors scaled to unit
    rows.  Plain FP64 Davidson against FP32 matvecs with FP64 refinement,
    from the same guess to the same residual tolerance."""
    import t3ns_amd as T
    from t3ns_amd.environment import Environment, normalised_rows
    t0 = time.perf_counter()
    env = Environment(ctx, seed=1, site_tensor=normalised_rows)
    d, o, s = env.step_operators(step)
    t_env = time.perf_counter() - t0
    h64 = T.Heff(d, o, s, T.FP64, keepalive=(step, env))
    h32 = T.Heff(d, o, s, T.FP32, keepalive=(step, env))
    print("# solve: operators in %.1f s, flops/matvec %.4g" % (t_env, h64.stats()["flops"]),
          file=sys.stderr, flush=True)
    fl = h64.stats()["flops"]
    rates = {}
    for name, h in (("f64", h64), ("f32", h32)):
        ms = _time_matvec(h)
        rates[name] = {"ms_per_matvec": ms, "tflops": fl / (ms * 1e-3) / 1e12,
                       "roofline_frac": fl / (ms * 1e-3) / 1e12 / PEAK_TFLOPS[name]}
    print("# solve matvec rates %s" % rates, file=sys.stderr, flush=True)
    guess = np.random.default_rng(2).standard_normal(h64.size)
    diag = h64.diagonal()
    t0 = time.perf_counter()
    e64, _, its, rc = h64.davidson(guess, tol=tol, max_its=400, diagonal=diag)
    t64 = time.perf_counter() - t0
    print("# solve FP64: E %.12f its %d (%.1f s)" % (e64, its, t64), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    em, _, its_lo, its_hi, rcm = T.davidson_refined(h32, h64, guess, tol=tol, switch_tol=tol,
                                                    max_its=400, diagonal=diag)
    tm = time.perf_counter() - t0
    print("# solve mixed: E %.12f its %d+%d (%.1f s)" % (em, its_lo, its_hi, tm),
          file=sys.stderr, flush=True)
    return {"operators": "device renormalisation chain from the vacuum, seeded site tensors "
                         "with unit rows", "operators_s": t_env,
            "flops_per_matvec": h64.stats()["flops"], "tol": tol, "matvec": rates,
            "fp64": {"energy": e64, "its": its, "status": rc, "seconds": t64},
            "mixed": {"energy": em, "its_f32": its_lo, "its_f64": its_hi, "status": rcm,
                      "seconds": tm},
            "abs_dE": abs(em - e64), "speedup_vs_fp64": t64 / tm}


def sweep_phases(dm, solver):
    """Seconds per phase of a sweep run under the reference's timer names
    (src/optimize_network.c:43-58): the sweep loop's and the solver's."""
    ph = dict(dm.timers)
    ph.update(solver.timers)
    return {k: round(v, 4) for k, v in sorted(ph.items())}


def measure_sweep(sweeps=4, e_conv=1e-10):
    """Two-site DMRG of the NN_HUBBARD(t = 1, U = 4) 10-site chain
    (makenetwork.py 10 DMRG), 10 electrons (U1 x U1: 5 up, 5 down), D = 64,
    up to 4 sweeps with E_CONV = 1e-10, the reference's default Davidson
    settings (tolerance 1e-5, 100 iterations); start: every block entry
    N(0, 1) from PCG64 seed 0 in site then block order, canonicalised by QR
    toward the first step.  Every step on the device (makesiteTensor,
    renormalisation, Heff + Davidson, split), FP64, and FP32 matvecs with
    FP64 refinement on every step (the counts of FP32 and FP64 steps and
    iterations are reported)."""
    import t3ns_amd as T
    from t3ns_amd.sweep import Sweeper, DeviceSolver
    with open(os.path.join(ROOT, "tests", "golden", "networks", "10_DMRG.netw")) as f:
        netw = f.read()
    out = {"workload": "nn_hubbard_10_chain_D64", "max_sweeps": sweeps, "e_conv": e_conv,
           "davidson_tol": 1e-5, "davidson_max_its": 100, "start": "N(0,1) seed 0, QR canonical"}
    for prec in ("f64", "mixed"):
        ctx = T.Context(netw, "NN_HUBBARD", [1.0, 4.0], ["U1", "U1"], [5, 5], 64)
        t0 = time.perf_counter()
        solver = DeviceSolver(prec, max_its=100, mixed_min_size=0)
        dm = Sweeper(ctx, 64, seed=0, solver=solver, tol=1e-5, start="normal")
        e = dm.run(sweeps, e_conv=e_conv)
        t = time.perf_counter() - t0
        out[prec] = {"energy": e, "sweeps": len(dm.energies), "seconds": t,
                     "seconds_per_sweep": t / len(dm.energies), "sweep_energies": dm.energies,
                     "counts": dict(solver.counts), "phases": sweep_phases(dm, solver)}
        print("# sweep %s: E %.12f in %d sweeps, %.2f s %s" % (prec, e, len(dm.energies), t,
                                                            solver.counts),
              file=sys.stderr, flush=True)
    out["abs_dE"] = abs(out["mixed"]["energy"] - out["f64"]["energy"])
    out["lif"] = measure_lif_sweep()
    return out


def measure_lif_sweep():
    """The reference's LiF regression run (tests/test3.c.in, singlet): 25
    orbitals, 6 electrons, Z2 x U1 x SU2 x C2v, lif_T3NS tree, two-site
    steps, the scheme of tests/test3.c.in:81-84 (minD = maxD = 100, 4
    Davidson iterations per step, at most 2 sweeps, E_CONV 1e-4; then 140,
    100 iterations, at most 10 sweeps, E_CONV 1e-6; truncation 1e-4,
    Davidson tolerance 1e-6); reference energy -107.1152169273.  The mixed
    leg runs FP32 matvecs with FP64 refinement on every step."""
    import gzip
    import t3ns_amd as T
    from t3ns_amd.sweep import DeviceSolver, Regime, Sweeper
    gold = os.path.join(ROOT, "tests", "golden")
    with gzip.open(os.path.join(gold, "fcidumps", "LiF_3.05.FCIDUMP.gz"), "rt") as f:
        fci = f.read()
    with open(os.path.join(gold, "ref_networks", "lif_T3NS.netw")) as f:
        netw = f.read()
    out = {"workload": "lif_t3ns_test3_singlet", "reference_energy": -107.1152169273}
    for prec in ("f64", "mixed"):
        ctx = T.Context(netw, "QC", [], ["U1", "SU2", "C2v"], [6, 0, 0], 100, fcidump=fci)
        t0 = time.perf_counter()
        solver = DeviceSolver(prec, mixed_min_size=0)
        dm = Sweeper(ctx, 100, seed=1, solver=solver)
        # the reference's scheme (tests/test3.c.in:81-84): 4 Davidson
        # iterations x at most 2 sweeps at D = 100, then 100 iterations x at
        # most 10 sweeps at D = 140
        e = dm.run_scheme([Regime(100, 100, 1e-4, 1e-6, 4, 2, 1e-4),
                           Regime(140, 140, 1e-4, 1e-6, 100, 10, 1e-6)])
        t = time.perf_counter() - t0
        out[prec] = {"energy": e, "abs_err_vs_reference": abs(e - out["reference_energy"]),
                     "sweeps": len(dm.energies), "seconds": t,
                     "seconds_per_sweep": t / len(dm.energies),
                     "regime_energies": dm.regime_energies, "counts": dict(solver.counts),
                     "phases": sweep_phases(dm, solver)}
        print("# lif sweep %s: E %.10f in %d sweeps, %.2f s" % (prec, e, len(dm.energies), t),
              file=sys.stderr, flush=True)
    return out


def measure_qc24_sweep(D=128, sweeps=1):
    """One two-site sweep of the QC-24 system (the matvec workload's
    Hamiltonian: synthetic FCIDUMP seed 1, 24 orbitals and electrons, ERI
    rank 48, U1 x SU2, makenetwork.py 24 tree) at bond dimension D from the
    seeded canonical start, FP64 and mixed: seconds, per-phase seconds,
    energies."""
    import t3ns_amd as T
    from t3ns_amd import synthetic
    from t3ns_amd.sweep import DeviceSolver, Sweeper
    fci = synthetic.qc_workload(24, 24, seed=1, rank=48)
    with open(NETWORK) as f:
        netw = f.read()
    out = {"workload": "qc24_t3ns_D%d" % D, "sweeps": sweeps, "davidson_tol": 1e-5}
    for prec in ("f64", "mixed"):
        ctx = T.Context(netw, "QC", [], ["U1", "SU2"], [24, 0], D, fcidump=fci)
        t0 = time.perf_counter()
        solver = DeviceSolver(prec, max_its=100, mixed_min_size=0)
        dm = Sweeper(ctx, D, seed=1, solver=solver, tol=1e-5)
        e = dm.run(sweeps)
        t = time.perf_counter() - t0
        out[prec] = {"energy": e, "seconds": t, "seconds_per_sweep": t / len(dm.energies),
                     "counts": dict(solver.counts), "phases": sweep_phases(dm, solver)}
        print("# qc24 sweep %s: E %.10f, %.2f s %s" % (prec, e, t, out[prec]["phases"]),
              file=sys.stderr, flush=True)
    out["abs_dE"] = abs(out["mixed"]["energy"] - out["f64"]["energy"])
    return out


def measure_serving(n, rate, slots=8):
    """The served-solve stream (NN_HUBBARD chains and trees, L ~ U{8..32},
    U ~ U[1, 8], D ~ U{64, 128, 256, 512}, half filling, U1 x U1, 4 sweeps
    each, Poisson arrivals at `rate`/s, seed 3), its first n requests:
    served with continuous admission (slots, HBM budget; each admitted
    solve on its own host thread and streams) against the same
    solves run one at a time in FP64 in arrival order (a FIFO queue under
    the same arrivals).  Reports solves/s, p50/p99 latency of both and the
    largest energy difference between a served result and the same solve
    alone."""
    from t3ns_amd import serve
    reqs = serve.hubbard_stream(n, rate, seed=3)
    t0 = time.perf_counter()
    recs = serve.Server(slots=slots).run(reqs)
    t_served = time.perf_counter() - t0
    served = serve.latency_summary(recs)
    print("# serve: %d solves in %.1f s, p50 %.2f s p99 %.2f s" %
          (n, t_served, served["p50_s"], served["p99_s"]), file=sys.stderr, flush=True)
    # one at a time, plain FP64, arrival order: service times, then the
    # FIFO schedule under the same arrivals
    service, energies = [], []
    for r in reqs:
        s0 = time.perf_counter()
        e, _ = serve.solve(r)
        service.append(time.perf_counter() - s0)
        energies.append(e)
    clock, lat = 0.0, []
    for r, sv in zip(reqs, service):
        clock = max(clock, r.arrival) + sv
        lat.append(clock - r.arrival)
    span = clock - reqs[0].arrival
    alone = {"solves": n, "solves_per_s": n / span, "p50_s": float(np.percentile(lat, 50)),
             "p99_s": float(np.percentile(lat, 99)), "makespan_s": span,
             "service_s_total": float(sum(service))}
    print("# serve alone: %.1f s of service, p50 %.2f s p99 %.2f s" %
          (sum(service), alone["p50_s"], alone["p99_s"]), file=sys.stderr, flush=True)
    dE = max(abs(rec["energy"] - e) for rec, e in zip(recs, energies))
    return {"workload": "nn_hubbard_stream_seed3", "requests": n, "rate_per_s": rate,
            "slots": slots, "workers": "one host thread per slot", "precision": "f64", "served": served, "one_at_a_time_f64": alone,
            "speedup_solves_per_s": served["solves_per_s"] / alone["solves_per_s"],
            "max_abs_dE_served_vs_alone": dE,
            "requests_LD": [[r.geometry, r.L, r.D] for r in reqs]}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    import t3ns_amd as T

    if args.copy_calibration:
        n = (256 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8}))
        return

    wl = WORKLOADS[args.workload]
    sites = [int(v) for v in args.sites.split(",")] if args.sites else list(wl["sites"])
    if args.bond_dim is None:
        args.bond_dim = wl["bond_dim"]
    if args.workload != "qc24":  # the other legs belong to the default workload
        args.no_solve = args.no_sweep = args.no_rows = args.no_serve = args.no_batch = True
    t0 = time.perf_counter()
    ctx, step = build_step(sites, args.bond_dim, args.workload)
    t_setup = time.perf_counter() - t0
    print("# setup %.1f s, vector size %d" % (t_setup, step.size), file=sys.stderr, flush=True)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(step.size)
    x /= np.linalg.norm(x)

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    dx = torch.from_numpy(x).to(dev)
    results = {}
    ys = {}
    heffs = {}
    for prec in args.precisions.split(","):
        t0 = time.perf_counter()
        h = T.Heff.from_step(step, {"f64": T.FP64, "f32": T.FP32, "bf16": T.BF16}[prec])
        t_init = time.perf_counter() - t0
        heffs[prec] = h
        print("# %s: Heff init %.1f s" % (prec, t_init), file=sys.stderr, flush=True)
        tune = None
        if args.autotune_trials > 0:
            t0 = time.perf_counter()
            tune = h.autotune(max_trials=args.autotune_trials)
            tune["seconds"] = time.perf_counter() - t0
            print("# %s: autotune %s" % (prec, tune), file=sys.stderr, flush=True)
        elif args.wpc:
            h.set_tuning(wpc=args.wpc)
        dy = torch.empty_like(dx)
        torch.cuda.synchronize()
        for _ in range(args.warmup):
            w0 = time.perf_counter()
            h.matvec_device(dx.data_ptr(), dy.data_ptr(), stream.cuda_stream)
            stream.synchronize()
            print("# %s: warmup matvec %.3f s" % (prec, time.perf_counter() - w0),
                  file=sys.stderr, flush=True)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            h.matvec_device(dx.data_ptr(), dy.data_ptr(), stream.cuda_stream)
        ev1.record(stream)
        stream.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        ms = ev0.elapsed_time(ev1) / args.steps
        st = h.stats()
        ks = h.kernel_stats()
        print("# %s: %s" % (prec, " ".join("%s=%.4g" % kv for kv in {**st, **ks}.items())),
              file=sys.stderr, flush=True)
        ys[prec] = dy.cpu().numpy()
        results[prec] = {"ms_per_matvec": ms, "matvec_per_s": 1000.0 / ms,
                         "wall_s": wall, "init_s": t_init, "flops": st["flops"],
                         "tflops": st["flops"] / (ms * 1e-3) / 1e12, "tuning": tune}
        print("# %s: %.3f ms/matvec, %.2f TFLOP/s (init %.1f s)" %
              (prec, ms, results[prec]["tflops"], t_init), file=sys.stderr, flush=True)

    main_prec = "f64" if "f64" in results else list(results)[0]
    r = results[main_prec]
    flops = r["flops"]
    out = {
        "metric": ("Heff matvec/s (QC-24 T3NS, D=512, 3-site branching step)" if args.workload == "qc24"
                   else "Heff matvec/s (QC-50 T3NS, D=2048, 2-site branching step)"),
        "value": r["matvec_per_s"],
        "unit": "matvec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_matvec"],
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": main_prec,
        "data": "synthetic (FCIDUMP seed %d; N(0,1) operator blocks and input vector)" % wl["seed"],
        "config": {"workload": "%s_t3ns_D%d_sites%s" % (args.workload, args.bond_dim,
                                                         "-".join(map(str, sites))),
                   "norb": wl["norb"], "nelec": wl["norb"], "symmetry": "U1xSU2",
                   "bond_dim": args.bond_dim,
                   "sites": sites, "vector_size": step.size,
                   "flops_per_matvec": flops, "setup_s": t_setup},
        "roofline": {"bound": "mfma", "achieved": r["tflops"], "peak": PEAK_TFLOPS[main_prec],
                     "unit": "TFLOP/s", "frac": r["tflops"] / PEAK_TFLOPS[main_prec],
                     "traffic": profiled_traffic(args.workload, sites, args.bond_dim) if main_prec == "f64"
                     else None},
    }
    if "f32" in results and "f64" in results:
        err = float(np.linalg.norm(ys["f32"] - ys["f64"]) / np.linalg.norm(ys["f64"]))
        out["mixed"] = {"dtype": "f32", "matvec_per_s": results["f32"]["matvec_per_s"],
                        "ms_per_matvec": results["f32"]["ms_per_matvec"],
                        "speedup_vs_f64": results["f32"]["matvec_per_s"] / r["matvec_per_s"],
                        "rel_err_vs_f64": err,
                        "roofline_frac": results["f32"]["tflops"] / PEAK_TFLOPS["f32"]}
    if "bf16" in results and "f64" in results:
        err = float(np.linalg.norm(ys["bf16"] - ys["f64"]) / np.linalg.norm(ys["f64"]))
        out.setdefault("mixed", {})["bf16"] = {
            "storage": "bf16 operator blocks, per-block power-of-two scales",
            "arithmetic": "f32 matrix cores, f32 accumulation",
            "matvec_per_s": results["bf16"]["matvec_per_s"],
            "ms_per_matvec": results["bf16"]["ms_per_matvec"],
            "speedup_vs_f64": results["bf16"]["matvec_per_s"] / r["matvec_per_s"],
            "rel_err_vs_f64": err,
            "roofline_frac": results["bf16"]["tflops"] / PEAK_TFLOPS["bf16"]}
    out["autotune"] = {p: results[p]["tuning"] for p in results}
    if not args.no_solve:
        out["solve"] = measure_solve(ctx, step, args.solve_tol)
    if not args.no_batch:
        out["batch"] = measure_batch()
    if not args.no_sweep:
        out["sweep"] = measure_sweep()
        if args.qc24_sweep:
            out["sweep"]["qc24"] = measure_qc24_sweep(args.qc24_sweep)
    if not args.no_serve:
        out["serve"] = measure_serving(args.serve_requests, args.serve_rate)
    if not args.no_rows:
        t0 = time.perf_counter()
        out["rows"] = measure_rows(step, sites, args.bond_dim)
        print("# rows measured in %.1f s" % (time.perf_counter() - t0), file=sys.stderr, flush=True)
    if args.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(step, x, args.cpu_seconds, flops)
        if not args.no_sweep:
            out["cpu_baseline"]["sweep"] = cpu_sweep_baseline()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
