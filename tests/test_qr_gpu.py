"""qr_step (src/siteTensor_decompose.c:1587-1630) on the device against the
oracle restatement: the same new bond dims and R spectrum (1e-12), Q and
the centre times R equal up to the sign of each bond vector (Householder
signs follow LAPACK's convention in both, so the signs agree unless a pivot
is zero)."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from qr_util import blocks, sector_columns, shared_bond
from test_qr_oracle import QR_CASES, check_qr, qcase_id, qcontext


@pytest.mark.gpu
@pytest.mark.parametrize("case", QR_CASES, ids=qcase_id)
def test_qr_matches_oracle(gpu, case):
    ctx = qcontext(case)
    for sa, sc in case[5]:
        print("case", qcase_id(case), "sites", sa, sc, flush=True)
        d = ctx.qr_problem(sa, sc, seed=3)
        want = O.qr_step(d)
        got, (flops, nbytes, ms) = T.qr_step(d)
        for key in ("Q_beginblock", "B_beginblock", "dims", "nsing"):
            assert np.array_equal(got[key], want[key]), key
        np.testing.assert_allclose(got["sing"], want["sing"], rtol=0,
                                   atol=1e-12 * max(1.0, float(np.max(want["sing"], initial=0))))
        assert got["entropy"] == pytest.approx(want["entropy"], rel=1e-10, abs=1e-12)
        la, lc = shared_bond(d)
        gq = sector_columns(blocks(d.A_nrblocks, d.A_qnumbers, got["Q_beginblock"], got["Q_tel"], d.ssA,
                                   {la: got["dims"]}), la)
        wq = sector_columns(blocks(d.A_nrblocks, d.A_qnumbers, want["Q_beginblock"], want["Q_tel"], d.ssA,
                                   {la: want["dims"]}), la)
        for s, w in wq.items():
            sg = np.sign(np.sum(gq[s] * w, axis=0))
            assert np.all(sg != 0)
            np.testing.assert_allclose(gq[s] * sg[None, :], w, rtol=0, atol=1e-10)
        # the device result is itself a QR of A
        check_qr(d, got)
        assert flops > 0 and ms >= 0
