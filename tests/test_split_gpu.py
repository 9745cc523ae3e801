"""The SVD split (split_of_site, src/siteTensor_decompose.c:1403-1444) on
the device against the oracle restatement: the same singular values
(1e-12 relative), the same truncation and block structure, and the same
kept vectors up to the sign of each singular pair (the sign of an SVD
vector is arbitrary; LAPACK's choice is not reproduced)."""
import os

import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from conftest import NETW
from test_split_oracle import CASES, _cid, _ids, _step, splittable


def _vectors(d, r):
    """Per kept sector (new index): V rows (k x columns) and U columns
    (rows x k), blocks concatenated in result order."""
    sv = d.siteV
    incl = [i for i in range(d.nrsites) if i != sv]
    bv = [d.ss[sv][k].bond for k in range(3)]
    for i in incl:
        for j in range(3):
            if d.ss[i][j].bond in bv:
                cs, cb, vb = i, j, bv.index(d.ss[i][j].bond)
    nb = len(r["bond_dims"])
    V, U = {}, {}
    for b in range(len(r["V_beginblock"]) - 1):
        ns = [d.ss[sv][k].nrSecs if k != vb else nb for k in range(3)]
        ids = _ids(int(r["V_qnumbers"][b]), ns)
        s, k = ids[vb], r["bond_dims"][ids[vb]]
        blk = r["V_tel"][r["V_beginblock"][b]:r["V_beginblock"][b + 1]]
        v0 = int(np.prod([d.ss[sv][kk].dims[ids[kk]] for kk in range(vb)])) if vb else 1
        m = np.moveaxis(blk.reshape((v0, k, -1), order="F"), 1, 0).reshape((k, -1), order="F")
        V.setdefault(s, []).append(m)
    nu = r["U_nrsites"]
    uq = r["U_qnumbers"].reshape(-1, nu)
    ci = incl.index(cs)
    for b in range(len(r["U_beginblock"]) - 1):
        ns = [d.ss[cs][k].nrSecs if k != cb else nb for k in range(3)]
        ids = _ids(int(uq[b][ci]), ns)
        s, k = ids[cb], r["bond_dims"][ids[cb]]
        blk = r["U_tel"][r["U_beginblock"][b]:r["U_beginblock"][b + 1]]
        t0 = 1
        for j, i in enumerate(incl):
            nsi = ns if i == cs else [d.ss[i][kk].nrSecs for kk in range(3)]
            di = [d.ss[i][kk].dims[x] if not (i == cs and kk == cb) else 1
                  for kk, x in enumerate(_ids(int(uq[b][j]), nsi))]
            if i < cs:
                t0 *= int(np.prod(di))
            elif i == cs:
                t0 *= int(np.prod(di[:cb])) if cb else 1
        m = np.moveaxis(blk.reshape((t0, k, -1), order="F"), 1, 2).reshape((-1, k), order="F")
        U.setdefault(s, []).append(m)
    return ({s: np.hstack(v) for s, v in V.items()}, {s: np.vstack(u) for s, u in U.items()})


def _compare(d):
    want = O.split_site(d)
    got, (flops, nbytes, ms) = T.split_site(d)
    for key in ("nsing", "kept", "bond_dims", "bond_irreps", "U_qnumbers", "U_beginblock",
                "V_qnumbers", "V_beginblock"):
        assert np.array_equal(got[key], want[key]), key
    scale = max(1.0, float(np.max(want["sing"]))) if want["sing"].size else 1.0
    np.testing.assert_allclose(got["sing"], want["sing"], rtol=0, atol=1e-12 * scale)
    assert got["norm_U"] == pytest.approx(want["norm_U"], rel=1e-12)
    for i in range(2):
        assert got["norm"][i] == pytest.approx(want["norm"][i], rel=1e-12)
        assert got["entropy"][i] == pytest.approx(want["entropy"][i], rel=1e-10, abs=1e-12)
    gv, gu = _vectors(d, got)
    wv, wu = _vectors(d, want)
    offs = np.concatenate([[0], np.cumsum(want["nsing"])])
    for s in wv:
        sg = np.sign(np.sum(gv[s] * wv[s], axis=1))
        assert np.all(sg != 0)
        np.testing.assert_allclose(gv[s] * sg[:, None], wv[s], rtol=0, atol=1e-9)
        np.testing.assert_allclose(gu[s] * sg[None, :], wu[s], rtol=0, atol=1e-9)
    return got


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_split_matches_oracle(gpu, case):
    ctx, st = _step(case)
    d0, o, s = st.descriptors()
    for sv in splittable(st, s.nrsites):
        for maxD in (case[-2], 100000):
            print("case", _cid(case), "site", sv, "maxD", maxD, flush=True)
            _compare(st.split_problem(sv, maxD=maxD))


@pytest.mark.gpu
@pytest.mark.parametrize("sv", [0, 2])
def test_split_qc24_d512(gpu, sv):
    from t3ns_amd import synthetic
    print("case qc24 D512 [18, 31, 17] site", sv, flush=True)
    fci = synthetic.qc_workload(24, 24, seed=1, rank=48)
    with open(os.path.join(NETW, "24_T3NS.netw")) as f:
        ctx = T.Context(f.read(), "QC", [], ["U1", "SU2"], [24, 0], 512, fcidump=fci)
    st = T.Step(ctx, [18, 31, 17], seed=1)
    _compare(st.split_problem(sv, maxD=512))


@pytest.mark.gpu
@pytest.mark.parametrize("case", [c for c in CASES if c[0] == "qc"][:4], ids=_cid)
def test_split_null_singular_values_match_oracle(gpu, case):
    """Kept null singular values: the device completes V's rows as the
    oracle does (orthonormal rows, checked on the oracle in
    test_split_oracle.py)."""
    from test_split_oracle import keep_first_block
    ctx, st = _step(case)
    d0, o, s = st.descriptors()
    for sv in splittable(st, s.nrsites):
        print("case", _cid(case), "site", sv, "null values", flush=True)
        d = st.split_problem(sv, maxD=100000)
        keep_first_block(d)
        _compare(d)


@pytest.mark.gpu
@pytest.mark.parametrize("case", [CASES[0], CASES[-1]], ids=_cid)
def test_split_jacobi_paths_agree(gpu, case, monkeypatch):
    """The one-workgroup Jacobi path for small sector matrices sums in the
    per-round path's order: both give the same bits (the QC-24 D = 512
    split has sectors over the small path's limit, so both paths also run
    together there)."""
    ctx, st = _step(case)
    d0, o, s = st.descriptors()
    for sv in splittable(st, s.nrsites):
        print("case", _cid(case), "site", sv, flush=True)
        d = st.split_problem(sv, maxD=case[-2])
        monkeypatch.delenv("T3NS_JACOBI_SMALL_COLS", raising=False)
        a, _ = T.split_site(d)
        monkeypatch.setenv("T3NS_JACOBI_SMALL_COLS", "0")
        b, _ = T.split_site(d)
        monkeypatch.delenv("T3NS_JACOBI_SMALL_COLS")
        assert a.keys() == b.keys()
        for key in a:
            assert np.array_equal(np.asarray(a[key]), np.asarray(b[key])), key
