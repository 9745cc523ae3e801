"""The large single-solve workload: synthetic FCIDUMP with 50 orbitals and
50 electrons (t3ns_amd.synthetic, seed 4, ERI rank 100), U1 x SU2, the T3NS
tree of makenetwork.py 50 (tests/golden/networks/50_T3NS.netw), bond
dimension 2048.  Two-site step 6 of the sweep (branching site 71 and
physical site 43, whose outer bond carries 2066 reduced states -- the
bookkeeper's D = 2048 cap) with seeded operators: the device matvec
against the oracle on sampled output blocks (the whole matvec is ~1e12
flops), FP64 1e-12 relative per block; the FP32 matvec within 1e-5 of the
FP64 one."""
import os

import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from t3ns_amd import synthetic

NETW = os.path.join(os.path.dirname(__file__), "golden", "networks", "50_T3NS.netw")


@pytest.mark.gpu
def test_qc50_D2048_step_sampled_blocks(gpu):
    fci = synthetic.qc_workload(50, 50, seed=4, rank=100)
    with open(NETW) as f:
        ctx = T.Context(f.read(), "QC", [], ["U1", "SU2"], [50, 0], 2048, fcidump=fci)
    sites = ctx.sweep_steps(2)[6]
    print("qc50 step", sites, flush=True)
    assert sorted(sites) == [43, 71]
    bonds = [b for site in sites for b in ctx.site_bonds(site)[0] if b >= 0]
    assert max(sum(ctx.bond_dims(b)) for b in bonds) > 2000
    st = T.Step(ctx, sites, seed=1)
    d, o, s = st.descriptors()
    h = T.Heff(d, o, s, T.FP64, keepalive=st)
    x = np.random.default_rng(4).standard_normal(h.size)
    y = h.matvec(x)
    h32 = T.Heff(d, o, s, T.FP32, keepalive=st)
    y32 = h32.matvec(x)
    assert np.linalg.norm(y32 - y) / np.linalg.norm(y) < 1e-5
    oh = O.OracleHeff(d, o, s)
    bf = oh.block_flops()
    sz = T.plan_block_dims(d, o, s).prod(axis=1)
    bb = np.ctypeslib.as_array(s.beginblock, shape=(s.nrblocks + 1,))
    live = np.where(bf > 0)[0]
    cheap = live[bf[live] <= 2e10]
    picks = sorted({int(cheap[np.argmax(sz[cheap])]),
                    int(live[np.argsort(bf[live])[len(live) // 2]]),
                    int(live[np.argmin(sz[live])])})
    print("blocks", picks, "sizes", sz[picks], "flops", bf[picks], flush=True)
    out = np.zeros(oh.size)
    for b in picks:
        oh.matvec_blocks(x, out, b, b + 1, 1)
    for b in picks:
        r = slice(int(bb[b]), int(bb[b + 1]))
        err = np.linalg.norm(y[r] - out[r]) / np.linalg.norm(out[r])
        assert err < 1e-12, (b, err)
    oh.close()
    h32.close()
    h.close()
    ctx.close()
