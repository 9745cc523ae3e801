"""Pins the oracle's Heff diagonal (restatement of make_diagonal,
src/Heff.c:1064-1108, 1268-1312) on the matvec itself: diag(H)[i] must equal
<e_i, H e_i> with H the oracle's matvec (the rule the reference checks in
its debug build, src/Heff.c:80-105).  The device diagonal is compared with
the oracle diagonal elsewhere, so this pins it transitively.

Objects of up to 1000 states are checked on every index, larger ones on 200
seeded indices, the QC-24 T3NS three-site step (D = 128) on 20.  A probe
<e_i, H e_i> computes only the output block holding i (matvec_blocks), so
it costs one block's share of a matvec."""
import os

import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, QC_FCIDUMP, case_id
from conftest import ROOT, netw_text

TOL = 1e-12
DENSE_MAX = 1000
NSAMPLE = 200


def _fcidump():
    with open(os.path.join(ROOT, "tests", "golden", QC_FCIDUMP)) as f:
        return f.read()


def _probe(oh, bb, idx):
    """<e_i, H e_i> for every i in idx, computing only the output block of
    i (bb: block start offsets)."""
    size = oh.size
    out = np.zeros(size)
    e = np.zeros(size)
    vals = []
    for i in idx:
        b = int(np.searchsorted(bb, i, side="right") - 1)
        e[i] = 1.0
        oh.matvec_blocks(e, out, b, b + 1, 1)
        vals.append(out[i])
        e[i] = 0.0
    return np.array(vals)


def _begins(site):
    return np.ctypeslib.as_array(site.beginblock, shape=(site.nrblocks + 1,))[:-1].copy()


def _check(oh, site, seed):
    n = oh.size
    dg = oh.diagonal()
    if n <= DENSE_MAX:
        idx = np.arange(n)
    else:
        idx = np.sort(np.random.default_rng(seed).choice(n, NSAMPLE, replace=False))
    ref = _probe(oh, _begins(site), idx)
    scale = max(np.max(np.abs(ref)), 1e-300)
    err = np.max(np.abs(dg[idx] - ref)) / scale
    assert err < TOL, err
    return len(idx)


@pytest.mark.parametrize("case", HEFF_CASES, ids=case_id)
def test_oracle_diagonal_is_unit_vector_matvec_hubbard(case):
    name, sy, ts, D, sites = case
    ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    st = T.Step(ctx, sites, seed=5)
    d, o, s = st.descriptors()
    oh = O.OracleHeff(d, o, s)
    _check(oh, s, 1)
    oh.close()


@pytest.mark.parametrize("case", QC_CASES, ids=case_id)
def test_oracle_diagonal_is_unit_vector_matvec_qc(case):
    name, sy, ts, D, sites = case
    ctx = T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=_fcidump())
    st = T.Step(ctx, sites, seed=7)
    d, o, s = st.descriptors()
    oh = O.OracleHeff(d, o, s)
    _check(oh, s, 2)
    oh.close()


def test_oracle_diagonal_qc24_branching_step():
    """The QC-24 T3NS benchmark step (sites 18, 31, 17) at D = 128: 20
    probed indices."""
    from t3ns_amd import synthetic
    fci = synthetic.qc_workload(24, 24, seed=1, rank=48)
    ctx = T.Context(netw_text("24_T3NS"), "QC", [], ["U1", "SU2"], [24, 0], 128, fcidump=fci)
    st = T.Step(ctx, [18, 31, 17], seed=1)
    d, o, s = st.descriptors()
    oh = O.OracleHeff(d, o, s)
    dg = oh.diagonal()
    idx = np.sort(np.random.default_rng(3).choice(oh.size, 20, replace=False))
    ref = _probe(oh, _begins(s), idx)
    err = np.max(np.abs(dg[idx] - ref)) / np.max(np.abs(ref))
    assert err < TOL, err
    oh.close()


def test_planted_diagonal_error_is_caught():
    """A diagonal off by one element's worth in one block fails the check
    (the comparison is not vacuous)."""
    name, sy, ts, D, sites = HEFF_CASES[3]
    ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    st = T.Step(ctx, sites, seed=5)
    d, o, s = st.descriptors()
    oh = O.OracleHeff(d, o, s)
    dg = oh.diagonal()
    idx = np.arange(oh.size)
    ref = _probe(oh, _begins(s), idx)
    bad = dg.copy()
    bad[oh.size // 2] *= 1.0 + 1e-6
    assert np.max(np.abs(dg - ref)) <= TOL * np.max(np.abs(ref))
    assert np.max(np.abs(bad - ref)) > TOL * np.max(np.abs(ref))
    oh.close()
