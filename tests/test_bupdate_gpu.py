"""update_rOperators_branching on the device against the oracle restatement
(src/rOperators_bUpdate.c:229-1101): same surviving blocks, elements within
1e-12 relative (the contractions run in another order)."""
import os

import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from conftest import NETW
from test_bupdate_oracle import BCASES, bcase_id, context, site_of


def _close(got, want, rtol):
    assert len(got) == len(want)
    for (gb, gt), (wb, wt) in zip(got, want):
        assert np.array_equal(gb, wb)
        if len(wt):
            assert np.max(np.abs(gt - wt)) <= rtol * max(1.0, np.max(np.abs(wt)))


@pytest.mark.gpu
@pytest.mark.parametrize("uc", [0, 1, 2])
@pytest.mark.parametrize("case", BCASES, ids=bcase_id)
def test_update_branching_matches_oracle(gpu, case, uc):
    print("case", bcase_id(case), "update case", uc, flush=True)
    ctx = context(case)
    d, o = ctx.bupdate_problem(site_of(ctx, case), uc, seed=21 + uc)
    got = T.update_branching(d, o)
    _close(got.blocks(), O.update_branching(d, o), 1e-12)
    flops, nbytes, ms = got.work()
    assert flops > 0 and nbytes > 0 and ms >= 0


@pytest.mark.gpu
@pytest.mark.parametrize("uc", [0, 2])
def test_update_branching_qc24(gpu, uc):
    """The branching site of the QC-24 benchmark object at D = 128 (blocks
    beyond one 64 x 64 tile, 11853 instructions)."""
    from t3ns_amd import synthetic
    print("case qc24 D128 site 31 update case", uc, flush=True)
    fci = synthetic.qc_workload(24, 24, seed=1, rank=48)
    with open(os.path.join(NETW, "24_T3NS.netw")) as f:
        ctx = T.Context(f.read(), "QC", [], ["U1", "SU2"], [24, 0], 128, fcidump=fci)
    d, o = ctx.bupdate_problem(31, uc, seed=3)
    _close(T.update_branching(d, o).blocks(), O.update_branching(d, o), 1e-12)
