"""Pins of the oracle's update_rOperators_physical restatement
(src/rOperators_pUpdate.c:17-387) that do not use the oracle's own formula.

The unity: appending a site to identity operators (oracle append restatement)
gives physical operators of which one is the identity.  Updating it with any
site tensor T must give, per sector s of the updated bond, X_s S X_s^T, where
X_s holds the rows of T in sector s over all (far bond, physical) columns and
S is 1 for left operators and the fermion parity (-1)^|phys| of the column
for right operators (the product of the adjoint and pUpdate signs,
src/symmetry_z2.c:62-103).  The expected value is assembled densely from
the tensor's blocks, independently of the block-sparse bookkeeping under test
(sector translation, usb_to_osb ranges, block lookup, transposes, dims,
empty-block removal).  SU(2) cases are not pinned this way: the SU(2) unity
carries multiplet factors (their prefactors are pinned against the compiled
reference sources in test_oracle_pins.py).

Norm preservation: without truncation and with a site tensor that is
orthogonal per sector, the update is a change of basis up to signs, so every
operator keeps its Frobenius norm -- for any operator values, every case
(SU(2) included, the SU(2) reduced update carries no extra factors) and both
directions.  A dropped or doubly counted block, a wrong tensor block or a
transposed operand changes the norm."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from conftest import netw_text
from renorm_util import (orthogonal_site_tensor, physical_rops, sector_rows, set_identity,
                         unity_ops)
from test_qc_heff import fcidump

ALL = [("hub",) + c for c in HEFF_CASES] + [("qc",) + c for c in QC_CASES]
CASES = [c for c in ALL if "SU2" not in c[2]]


def _cid(c):
    return c[0] + "-" + case_id(c[1:])


def _step(case):
    kind, name, sy, ts, D, sites = case
    if kind == "hub":
        ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    else:
        ctx = T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=fcidump())
    return ctx, T.Step(ctx, sites, seed=4)


@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_updated_unity(case):
    ctx, st = _step(case)
    d, o, s = st.descriptors()
    nchecked = 0
    for leg in range(3):
        if o[leg].nrhss <= 0 or not o[leg].P_operator:
            continue
        keep = []
        ad, old = st.append_problem(leg, seed=3)
        set_identity(old, ad.old_ss[0].nrSecs, keep)
        app = O.append_phys(ad, old)
        unity = unity_ops(ad, app)
        assert unity, "no identity among the appended operators"
        pd, _ = st.pupdate_problem(leg, seed=5, maxdim=4)
        pops = physical_rops(ad, app, keep)
        upd = O.update_physical(pd, pops)
        il = pd.is_left
        ub = 2 if il else 0
        nsu = pd.site_ss[ub].nrSecs
        dims = [pd.site_ss[ub].dims[i] for i in range(nsu)]
        for k in unity:
            bb, tel = upd[k]
            h = pops.hss_of_ops[k]
            b0 = pd.begin_blocks_of_hss[h]
            for b in range(len(bb) - 1):
                q = pd.qnumbers[b0 + b]
                bra, ket = q % nsu, (q // nsu) % nsu
                blk = tel[bb[b]:bb[b + 1]]
                if bra != ket:
                    assert len(blk) == 0
                    continue
                X, sg = sector_rows(pd, bra)
                want = (X * (1.0 if il else sg)) @ X.T
                if not np.any(np.abs(want) >= 1e-20):
                    assert len(blk) == 0
                    continue
                assert len(blk) == dims[bra] ** 2
                got = blk.reshape((dims[bra], dims[bra]), order="F")
                assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want)))
                nchecked += 1
    assert nchecked > 0


def _norms(blocks):
    return np.array([np.linalg.norm(t) for _, t in blocks])


@pytest.mark.parametrize("case", ALL, ids=_cid)
def test_update_preserves_norms(case):
    ctx, st = _step(case)
    d, o, s = st.descriptors()
    nlegs = 0
    for leg in range(3):
        if o[leg].nrhss <= 0 or not o[leg].P_operator:
            continue
        keep = []
        pd, pops = st.pupdate_problem(leg, seed=9, maxdim=0)
        orthogonal_site_tensor(pd, np.random.default_rng(leg), keep)
        before = []
        for k in range(pops.nrops):
            h = pops.hss_of_ops[k]
            nb = pops.begin_blocks_of_hss[h + 1] - pops.begin_blocks_of_hss[h]
            n = pops.beginblock[k][nb]
            before.append(np.linalg.norm(np.ctypeslib.as_array(pops.tel[k], shape=(max(n, 1),))[:n]))
        after = _norms(O.update_physical(pd, pops))
        before = np.array(before)
        assert np.any(before > 0)
        np.testing.assert_allclose(after, before, rtol=1e-12, atol=0)
        nlegs += 1
    assert nlegs > 0
