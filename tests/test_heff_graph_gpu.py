"""Matvec replay from captured HIP graphs (t3ns_heff_use_graphs): every
replay equals the eager matvec exactly (the same kernels in the same
order), across replays separated by unrelated eager work and
synchronisations, for alternating input/output arrays, FP64 and FP32; a
Davidson solve with graphs on gives the eager solve's energy and
iteration count."""
import time

import numpy as np
import pytest

import t3ns_amd as T
from cases import HEFF_CASES, case_id
from conftest import netw_text


def _step(case, seed=5):
    name, sy, ts, D, sites = case
    ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    return ctx, T.Step(ctx, sites, seed=seed)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [T.FP64, T.FP32], ids=["f64", "f32"])
@pytest.mark.parametrize("case", HEFF_CASES[3:6], ids=case_id)
def test_graph_replays_equal_eager(gpu, case, prec):
    import torch
    print("graph case", case_id(case), prec, flush=True)
    ctx, st = _step(case)
    h = T.Heff.from_step(st, prec)
    dev = torch.device("cuda:0")
    s = torch.cuda.Stream(device=dev)
    rng = np.random.default_rng(1)
    xs = [torch.from_numpy(rng.standard_normal(h.size)).to(dev) for _ in range(2)]
    ys = [torch.empty_like(xs[0]) for _ in range(2)]
    torch.cuda.synchronize()
    eager = []
    for x, y in zip(xs, ys):
        h.matvec_device(x.data_ptr(), y.data_ptr(), s.cuda_stream)
        s.synchronize()
        eager.append(y.clone())
    h.use_graphs(True)
    junk = torch.zeros(1 << 20, device=dev, dtype=torch.float64)
    for rep in range(6):
        k = rep % 2
        ys[k].fill_(np.nan)
        torch.cuda.synchronize()
        h.matvec_device(xs[k].data_ptr(), ys[k].data_ptr(), s.cuda_stream)
        s.synchronize()
        junk.add_(1.0)  # unrelated eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(ys[k], eager[k]), (rep, k)
    h.use_graphs(False)
    h.close()


@pytest.mark.gpu
def test_davidson_with_graphs_equals_eager(gpu):
    ctx, st = _step(HEFF_CASES[4])
    h = T.Heff.from_step(st, T.FP64)
    guess = np.random.default_rng(2).standard_normal(h.size)
    e0, x0, i0, r0 = h.davidson(guess, tol=1e-6, max_its=60)
    h.use_graphs(True)
    t0 = time.perf_counter()
    e1, x1, i1, r1 = h.davidson(guess, tol=1e-6, max_its=60)
    t1 = time.perf_counter() - t0
    h.use_graphs(False)
    t0 = time.perf_counter()
    h.davidson(guess, tol=1e-6, max_its=60)
    t2 = time.perf_counter() - t0
    print("davidson %d its: graphs %.3f s, eager %.3f s" % (i1, t1, t2))
    assert (e1, i1, r1) == (e0, i0, r0)
    assert np.array_equal(x1, x0)
