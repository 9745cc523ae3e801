"""rOperators_append_phys on the device against the oracle restatement
(src/rOperators_pUpdate.c:393-673, sum_unique_rOperators, kick_zero_blocks)
on the physical legs of Hubbard and QC steps: identical block structure
(the same all-zero blocks dropped) and elements within 1e-14 relative."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from conftest import netw_text
from test_qc_heff import fcidump

CASES = [("hub",) + c for c in HEFF_CASES] + [("qc",) + c for c in QC_CASES]


def _step(case):
    kind, name, sy, ts, D, sites = case
    if kind == "hub":
        ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    else:
        ctx = T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=fcidump())
    return ctx, T.Step(ctx, sites, seed=4)


def _cid(c):
    return c[0] + "-" + case_id(c[1:])


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_append_phys_matches_oracle(gpu, case):
    ctx, st = _step(case)
    print("case", _cid(case), flush=True)
    d, o, s = st.descriptors()
    nphys = 0
    for leg in range(3):
        if o[leg].nrhss <= 0 or not o[leg].P_operator:
            continue
        nphys += 1
        ad, old = st.append_problem(leg, seed=11 + leg)
        want = O.append_phys(ad, old)
        got = T.append_phys(ad, old).blocks()
        assert len(got) == len(want)
        for (gb, gt), (wb, wt) in zip(got, want):
            assert np.array_equal(gb, wb)
            if len(wt):
                assert np.max(np.abs(gt - wt)) <= 1e-14 * max(1.0, np.max(np.abs(wt)))
    assert nphys > 0
