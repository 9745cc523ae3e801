"""Heff matvec parity on the benchmark object itself: the 24-orbital
synthetic quantum-chemistry Hamiltonian (t3ns_amd.synthetic, seed 1) on the
T3NS network tests/golden/networks/24_T3NS.netw, bond dimension 512, sites
18-31-17.  The whole matvec (6e12 flops) is out of the oracle's reach, so a
sample of output blocks is compared: the largest block the oracle computes
in a few seconds (its input blocks and stage outputs are the ones that leave
LDS), a median one and a small one.  FP64, relative error 1e-12 per block."""
import os

import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from t3ns_amd import synthetic

NETW = os.path.join(os.path.dirname(__file__), "golden", "networks", "24_T3NS.netw")


@pytest.mark.gpu
def test_matvec_fullsize_sampled_blocks(gpu):
    fci = synthetic.qc_workload(24, 24, seed=1, rank=48)
    with open(NETW) as f:
        ctx = T.Context(f.read(), "QC", [], ["U1", "SU2"], [24, 0], 512, fcidump=fci)
    st = T.Step(ctx, [18, 31, 17], seed=1)
    d, o, s = st.descriptors()
    h = T.Heff(d, o, s, T.FP64, keepalive=st)
    assert h.kernel_stats()["pairs_global_x"] > 0  # the large-block paths run
    x = np.random.default_rng(2).standard_normal(h.size)
    y = h.matvec(x)
    oh = O.OracleHeff(d, o, s)
    bf = oh.block_flops()
    sz = T.plan_block_dims(d, o, s).prod(axis=1)
    bb = np.ctypeslib.as_array(s.beginblock, shape=(s.nrblocks + 1,))
    live = np.where(bf > 0)[0]
    cheap = live[bf[live] <= 1.3e10]
    picks = [int(cheap[np.argmax(sz[cheap])]),
             int(live[np.argsort(bf[live])[len(live) // 2]]),
             int(live[np.argmin(sz[live])])]
    print("blocks", picks, "sizes", sz[picks], "flops", bf[picks], flush=True)
    assert sz[picks[0]] > 4096
    out = np.zeros(oh.size)
    for b in picks:
        oh.matvec_blocks(x, out, b, b + 1, 1)
    for b in picks:
        r = slice(int(bb[b]), int(bb[b + 1]))
        err = np.linalg.norm(y[r] - out[r]) / np.linalg.norm(out[r])
        assert err < 1e-12, (b, err)
    h.close()
    ctx.close()
