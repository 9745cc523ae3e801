"""Physical renormalisation on the device against the oracle restatements on
the physical legs of Hubbard and QC steps: rOperators_append_phys
(src/rOperators_pUpdate.c:393-673, sum_unique_rOperators, kick_zero_blocks;
elements within 1e-14 relative) and update_rOperators_physical
(src/rOperators_pUpdate.c:17-387; within 1e-12 relative), both with
identical block structure (the same all-zero blocks dropped)."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from conftest import netw_text
from renorm_util import physical_rops
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


def _close(got, want, rtol):
    assert len(got) == len(want)
    for (gb, gt), (wb, wt) in zip(got, want):
        assert np.array_equal(gb, wb)
        if len(wt):
            assert np.max(np.abs(gt - wt)) <= rtol * max(1.0, np.max(np.abs(wt)))


@pytest.mark.gpu
@pytest.mark.parametrize("maxdim", [3, 0])
@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_update_physical_matches_oracle(gpu, case, maxdim):
    """update_rOperators_physical (src/rOperators_pUpdate.c:364-387) on the
    device: same surviving blocks as the oracle, elements within 1e-12
    relative (the two GEMMs per block pair sum in another order)."""
    ctx, st = _step(case)
    print("case", _cid(case), "maxdim", maxdim, flush=True)
    d, o, s = st.descriptors()
    nphys = 0
    for leg in range(3):
        if o[leg].nrhss <= 0 or not o[leg].P_operator:
            continue
        nphys += 1
        pd, pops = st.pupdate_problem(leg, seed=21 + leg, maxdim=maxdim)
        want = O.update_physical(pd, pops)
        got = T.update_physical(pd, pops)
        _close(got.blocks(), want, 1e-12)
        flops, nbytes, ms = got.work()
        assert flops > 0 and nbytes > 0 and ms >= 0
    assert nphys > 0


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_append_then_update_on_device(gpu, case):
    """The two physical renormalisation steps chained on the device (append
    output fed to the update) against the same chain in the oracle."""
    ctx, st = _step(case)
    print("case", _cid(case), flush=True)
    d, o, s = st.descriptors()
    for leg in range(3):
        if o[leg].nrhss <= 0 or not o[leg].P_operator:
            continue
        keep = []
        ad, old = st.append_problem(leg, seed=31 + leg)
        pd, _ = st.pupdate_problem(leg, seed=41 + leg, maxdim=5)
        dev_app = T.append_phys(ad, old)
        dev = T.update_physical(pd, dev_app)
        ora = O.update_physical(pd, physical_rops(ad, O.append_phys(ad, old), keep))
        _close(dev.blocks(), ora, 1e-12)


@pytest.mark.gpu
def test_update_physical_large_blocks(gpu):
    """Blocks larger than one 64 x 64 tile in every dimension (QC, N2 on the
    DMRG chain, sectors up to 150 wide)."""
    case = ("qc",) + QC_CASES[-1]
    ctx, st = _step(case)
    d, o, s = st.descriptors()
    for leg in range(3):
        if o[leg].nrhss <= 0 or not o[leg].P_operator:
            continue
        pd, pops = st.pupdate_problem(leg, seed=51 + leg, maxdim=150)
        print("leg", leg, "T blocks", pd.T_nrblocks, flush=True)
        want = O.update_physical(pd, pops)
        got = T.update_physical(pd, pops)
        _close(got.blocks(), want, 1e-12)
