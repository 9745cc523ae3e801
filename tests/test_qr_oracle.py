"""Pins of the oracle's qr_step restatement (src/siteTensor_decompose.c:
1587-1630) that do not use its own formulas: Q has orthonormal columns over
the bond in every sector, contracting Q with the centre times R equals
contracting A with the centre (QR = A), and R's singular values are those
of A's sector matrices (numpy's SVD)."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from conftest import netw_text
from qr_util import blocks, contract_pair, sector_columns, shared_bond
from test_qc_heff import fcidump

# (model, network, symmetries, target, D, site pairs (A, centre))
QR_CASES = [
    ("hub", "10_DMRG", ["U1", "U1"], [5, 5], 64, [(4, 5), (5, 4), (2, 3)]),
    ("hub", "10_T3NS", ["U1", "SU2"], [10, 0], 24, [(0, 10), (2, 10), (10, 2)]),
    ("qc", "ref_networks/10_T3NS", ["U1", "U1", "D2h"], [7, 7, 0], 64, [(7, 8), (3, 8), (9, 8)]),
    ("qc", "ref_networks/10_DMRG", ["U1", "SU2"], [14, 0], 64, [(4, 5), (6, 5)]),
]


def qcase_id(c):
    return "%s-%s-%s-D%d" % (c[0], c[1], "x".join(c[2]), c[4])


def qcontext(c):
    kind, name, sy, ts, D, _ = c
    if kind == "hub":
        return T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    return T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=fcidump())


def check_qr(d, r):
    la, lc = shared_bond(d)
    A = blocks(d.A_nrblocks, d.A_qnumbers, d.A_beginblock, d.A_tel, d.ssA)
    C = blocks(d.C_nrblocks, d.C_qnumbers, d.C_beginblock, d.C_tel, d.ssC)
    newdims = r["dims"]
    Q = blocks(d.A_nrblocks, d.A_qnumbers, r["Q_beginblock"], r["Q_tel"], d.ssA, {la: newdims})
    B = blocks(d.C_nrblocks, d.C_qnumbers, r["B_beginblock"], r["B_tel"], d.ssC, {lc: newdims})
    # orthonormal columns
    for s, m in sector_columns(Q, la).items():
        np.testing.assert_allclose(m.T @ m, np.eye(m.shape[1]), rtol=0, atol=1e-12)
    # A C = Q (R C)
    want, got = contract_pair(A, C, la, lc), contract_pair(Q, B, la, lc)
    for key, t in want.items():
        g = got.get(key)
        if g is None:
            assert np.max(np.abs(t)) < 1e-12
            continue
        np.testing.assert_allclose(g, t, rtol=0, atol=1e-11 * max(1.0, np.max(np.abs(t))))
    # R's spectrum = A's spectrum per sector
    cols = sector_columns(A, la)
    offs = np.concatenate([[0], np.cumsum(r["nsing"])])
    for s, m in cols.items():
        sv = np.linalg.svd(m, compute_uv=False)
        np.testing.assert_allclose(r["sing"][offs[s]:offs[s + 1]], sv[:r["nsing"][s]], rtol=0,
                                   atol=1e-12 * max(1.0, sv[0]))
    return len(want)


@pytest.mark.parametrize("case", QR_CASES, ids=qcase_id)
def test_qr_oracle_is_a_qr(case):
    ctx = qcontext(case)
    for sa, sc in case[5]:
        d = ctx.qr_problem(sa, sc, seed=3)
        assert check_qr(d, O.qr_step(d)) > 0
