"""Quantum-chemistry Heff on the reference's N2/STO-3G test inputs: the host
plan (from the QC operator and instruction generation) agrees exactly with
the oracle's restatement of src/Heff.c, and the device matvec and diagonal
agree with the oracle to 1e-12 (FP64)."""
import os

import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import QC_CASES, QC_FCIDUMP, case_id
from conftest import ROOT, netw_text
from plan_util import without_zero_couplings


def fcidump(name=QC_FCIDUMP):
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return f.read()


def _setup(case, seed=7):
    name, sy, ts, D, sites = case
    ctx = T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=fcidump())
    st = T.Step(ctx, sites, seed=seed)
    return ctx, st


@pytest.mark.parametrize("case", QC_CASES, ids=case_id)
def test_qc_plan_matches_oracle(case):
    ctx, st = _setup(case)
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


@pytest.mark.gpu
@pytest.mark.parametrize("case", QC_CASES, ids=case_id)
def test_qc_matvec_and_diagonal(gpu, case):
    ctx, st = _setup(case)
    print("case", case_id(case), flush=True)
    d, o, s = st.descriptors()
    h = T.Heff(d, o, s, T.FP64, keepalive=st)
    oh = O.OracleHeff(d, o, s)
    x = np.random.default_rng(2).standard_normal(h.size)
    y, yr = h.matvec(x), oh.matvec(x)
    assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-12
    dg, dr = h.diagonal(), oh.diagonal()
    np.testing.assert_allclose(dg, dr, rtol=1e-12, atol=1e-12 * np.max(np.abs(dr)))
