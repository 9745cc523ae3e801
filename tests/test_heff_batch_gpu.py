"""Batched matvecs (t3ns_heff_batch_*): the work units of several objects
share one launch per kernel; every object's result must equal its own
matvec bitwise (a unit's work does not depend on the units around it), for
any mix of objects, kernels (input block in LDS or global memory, stage
outputs in global scratch) and precisions, over repeated calls with
different subsets on one batch."""
import numpy as np
import pytest

import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from test_heff_gpu import _setup as hub_setup
from test_qc_heff import _setup as qc_setup


def _objects(prec, monkeypatch=None, placements=(None,)):
    out = []
    cases = [("hub", c) for c in HEFF_CASES[:6]] + [("qc", c) for c in QC_CASES[:5]]
    for k, (kind, case) in enumerate(cases):
        if kind == "hub":
            ctx, st, d, o, s = hub_setup(case)
        else:
            ctx, st = qc_setup(case)
            d, o, s = st.descriptors()
        pl = placements[k % len(placements)]
        if monkeypatch is not None:
            if pl is None:
                monkeypatch.delenv("T3NS_HEFF_PLACEMENT", raising=False)
            else:
                monkeypatch.setenv("T3NS_HEFF_PLACEMENT", pl)
        out.append((T.Heff(d, o, s, prec, keepalive=st), ctx))
    if monkeypatch is not None:
        monkeypatch.delenv("T3NS_HEFF_PLACEMENT", raising=False)
    return [h for h, _ in out], [c for _, c in out]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [T.FP64, T.FP32, T.BF16], ids=["f64", "f32", "bf16"])
def test_batch_equals_own_matvecs(gpu, prec):
    print("batch precision", prec, flush=True)
    hs, keep = _objects(prec)
    rng = np.random.default_rng(8)
    xs = [rng.standard_normal(h.size) for h in hs]
    own = [h.matvec(x) for h, x in zip(hs, xs)]
    b = T.HeffBatch()
    got = b.matvec(hs, xs)
    for i, (g, w) in enumerate(zip(got, own)):
        assert np.array_equal(g, w), i
    # other subsets on the same batch (buffers reused, lists rebuilt)
    for sub in ([0], [3, 1], list(range(len(hs)))[::-1], [2, 5, 7, 9]):
        got = b.matvec([hs[i] for i in sub], [xs[i] for i in sub])
        for j, i in enumerate(sub):
            assert np.array_equal(got[j], own[i]), (sub, i)
    b.close()


@pytest.mark.gpu
def test_batch_mixes_kernels(gpu, monkeypatch):
    """Objects whose units take other kernel paths (input blocks staged in
    LDS, stage outputs in global scratch; forced), batched with ordinary
    ones."""
    print("batch placements", flush=True)
    hs, keep = _objects(T.FP64, monkeypatch, placements=(None, "lds", "global"))
    rng = np.random.default_rng(9)
    xs = [rng.standard_normal(h.size) for h in hs]
    own = [h.matvec(x) for h, x in zip(hs, xs)]
    got = T.HeffBatch().matvec(hs, xs)
    for i, (g, w) in enumerate(zip(got, own)):
        assert np.array_equal(g, w), i


@pytest.mark.gpu
def test_batch_rejects_mixed_precisions(gpu):
    ctx, st, d, o, s = hub_setup(HEFF_CASES[0])
    h64 = T.Heff(d, o, s, T.FP64, keepalive=st)
    h32 = T.Heff(d, o, s, T.FP32, keepalive=st)
    with pytest.raises(RuntimeError):
        T.HeffBatch().matvec([h64, h32], [np.ones(h64.size), np.ones(h32.size)])
