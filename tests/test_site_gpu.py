"""makesiteTensor on the device against the oracle restatement
(src/siteTensor_init.c:609-800): all elements within 1e-12 relative to the
largest, zero blocks where a one-site block is missing."""
import os

import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from conftest import NETW, netw_text
from test_qc_heff import fcidump

CASES = [("hub",) + c for c in HEFF_CASES] + [("qc",) + c for c in QC_CASES]


def _cid(c):
    return c[0] + "-" + case_id(c[1:])


def _check(md):
    want = O.make_site_tensor(md)
    got, (flops, nbytes, ms) = T.make_site_tensor(md)
    assert got.shape == want.shape
    scale = max(1.0, float(np.max(np.abs(want))))
    assert np.max(np.abs(got - want)) <= 1e-12 * scale
    assert np.array_equal(got == 0, want == 0)
    assert flops > 0 and nbytes > 0 and ms >= 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_make_site_tensor_matches_oracle(gpu, case):
    kind, name, sy, ts, D, sites = case
    print("case", _cid(case), flush=True)
    if kind == "hub":
        ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    else:
        ctx = T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=fcidump())
    st = T.Step(ctx, sites, seed=4)
    _check(st.makesite_problem(seed=7))


@pytest.mark.gpu
def test_make_site_tensor_qc24_d512(gpu):
    """The three-site object of the QC-24 benchmark at D = 512 (sector dims
    beyond one 64 x 64 tile)."""
    from t3ns_amd import synthetic
    print("case qc24 D512 [18, 31, 17]", flush=True)
    fci = synthetic.qc_workload(24, 24, seed=1, rank=48)
    with open(os.path.join(NETW, "24_T3NS.netw")) as f:
        ctx = T.Context(f.read(), "QC", [], ["U1", "SU2"], [24, 0], 512, fcidump=fci)
    st = T.Step(ctx, [18, 31, 17], seed=1)
    _check(st.makesite_problem(seed=7))
