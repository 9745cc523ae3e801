"""The dispatch tables of the product (host plan builder) and of the oracle
restatement of src/Heff.c must agree exactly: same (new block, old block,
order, MPO group, operator blocks) entries in the same order, prefactors to
1e-15.  Runs on the CPU (no device needed)."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, case_id
from conftest import netw_text
from plan_util import without_zero_couplings


@pytest.mark.parametrize("case", HEFF_CASES, ids=case_id)
def test_plan_table_matches_oracle(case):
    name, sy, ts, D, sites = case
    ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    st = T.Step(ctx, sites, seed=3)
    d, o, s = st.descriptors()
    ints, prefs, flops = T.plan_table(d, o, s)
    oh = O.OracleHeff(d, o, s)
    oi, op, ndropped = without_zero_couplings(*oh.plan_table())
    assert ints.shape[0] > 0
    assert np.array_equal(ints, oi)
    np.testing.assert_allclose(prefs, op, rtol=0, atol=1e-15)
    if ndropped == 0:
        assert flops == oh.plan_stats()["flops"]
    else:
        assert flops < oh.plan_stats()["flops"]
