"""Pins of the oracle's split_of_site restatement
(src/siteTensor_decompose.c:587-1444) that do not use its own formulas:
what makes a result an SVD of the multisite tensor.

* Reconstruction: without truncation, U and V contracted over the cut bond
  (times the norm divided out of U) give back every block of A.
* V has orthonormal rows over the cut bond in every kept sector, U
  orthogonal columns whose norms are the singular values (over norm_U).
* Truncation (selectS, 1196-1289) keeps the maxD largest singular values
  overall, and the reported norm after truncation is their squared sum.
Together these fix the singular values, the kept vectors up to sign and
the block bookkeeping (layout of the sector matrices, U and V block order,
renumbered cut-bond sectors)."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from conftest import netw_text
from test_qc_heff import fcidump

CASES = [("hub",) + c for c in HEFF_CASES] + [("qc",) + c for c in QC_CASES]


def _cid(c):
    return c[0] + "-" + case_id(c[1:])


def _step(case):
    kind, name, sy, ts, D, sites = case
    if kind == "hub":
        ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    else:
        ctx = T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=fcidump())
    return ctx, T.Step(ctx, sites, seed=4)


def splittable(st, nsites):
    out = []
    for sv in range(nsites):
        try:
            O.split_site(st.split_problem(sv, maxD=1))
            out.append(sv)
        except RuntimeError:
            pass
    return out


def _ids(q, ns):
    return q % ns[0], (q // ns[0]) % ns[1], q // (ns[0] * ns[1])


class Split:
    """Block access to a split result over the descriptor's sectors."""

    def __init__(self, d, r):
        self.d, self.r = d, r
        n, sv = d.nrsites, d.siteV
        self.sv = sv
        self.incl = [i for i in range(n) if i != sv]
        bv = [d.ss[sv][k].bond for k in range(3)]
        for i in self.incl:
            for j in range(3):
                if d.ss[i][j].bond in bv:
                    self.cs, self.cb, self.vb = i, j, bv.index(d.ss[i][j].bond)
        cut = d.ss[sv][self.vb]
        self.cut_irr = [tuple(cut.irreps[s * 5 + k] for k in range(5)) for s in range(cut.nrSecs)]
        self.bond_irr = [tuple(x) for x in r["bond_irreps"]]
        self.newid = {s: (self.bond_irr.index(ir) if ir in self.bond_irr else -1)
                      for s, ir in enumerate(self.cut_irr)}

    def nsec(self, i):
        return [self.d.ss[i][k].nrSecs for k in range(3)]

    def dims(self, i, q, cutdim=None, renum=None):
        ns = self.nsec(i)
        ids = list(_ids(q, ns))
        out = []
        for k in range(3):
            is_cut = (i == self.sv and k == self.vb) or (i == self.cs and k == self.cb)
            if is_cut and cutdim is not None:
                out.append(cutdim)
            else:
                out.append(self.d.ss[i][k].dims[ids[k]])
        return out, ids

    def renumber(self, i, q):
        ns = self.nsec(i)
        ids = list(_ids(q, ns))
        k = self.vb if i == self.sv else self.cb
        s = ids[k]
        ids[k] = self.newid[s]
        nn = list(ns)
        nn[k] = len(self.bond_irr)
        return ids[0] + ids[1] * nn[0] + ids[2] * nn[0] * nn[1], s

    def ublock(self, tup):
        r = self.r
        nu = r["U_nrsites"]
        q = r["U_qnumbers"].reshape(-1, nu) if nu else r["U_qnumbers"]
        for b in range(len(r["U_beginblock"]) - 1):
            if tuple(q[b]) == tuple(tup):
                return r["U_tel"][r["U_beginblock"][b]:r["U_beginblock"][b + 1]]
        return None

    def vblock(self, q):
        r = self.r
        for b in range(len(r["V_beginblock"]) - 1):
            if r["V_qnumbers"][b] == q:
                return r["V_tel"][r["V_beginblock"][b]:r["V_beginblock"][b + 1]]
        return None


def _a_blocks(d):
    n = d.nrsites
    A = np.ctypeslib.as_array(d.tel, shape=(d.beginblock[d.nrblocks],))
    for b in range(d.nrblocks):
        yield [d.qnumbers[b * n + i] for i in range(n)], A[d.beginblock[b]:d.beginblock[b + 1]]


@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_split_reconstructs_and_is_an_svd(case):
    ctx, st = _step(case)
    d0, o, s = st.descriptors()
    svs = splittable(st, s.nrsites)
    assert svs
    for sv in svs:
        d = st.split_problem(sv, maxD=100000)
        r = O.split_site(d)
        assert int(r["kept"].sum()) == int(r["nsing"].sum()) or np.min(r["sing"]) < 1e-300
        check_split(d, r)


def keep_first_block(d):
    """Zeroes every block of the split problem's tensor but the first: most
    sector matrices become zero and their singular values null."""
    n = d.beginblock[d.nrblocks]
    A = np.ctypeslib.as_array(d.tel, shape=(n,))
    A[d.beginblock[1]:] = 0


@pytest.mark.parametrize("case", [c for c in CASES if c[0] == "qc"][:4], ids=_cid)
def test_split_null_singular_values_keep_v_orthonormal(case):
    """Kept null singular values (minD above the rank, as in the reference's
    LiF regimes with minD = maxD): V still has orthonormal rows per sector,
    as dgesdd's VT does (src/siteTensor_decompose.c:1060)."""
    ctx, st = _step(case)
    d0, o, s = st.descriptors()
    nnull = 0
    for sv in splittable(st, s.nrsites):
        d = st.split_problem(sv, maxD=100000)
        keep_first_block(d)
        r = O.split_site(d)
        assert int(r["kept"].sum()) == int(r["nsing"].sum())
        nnull += int(np.sum(r["sing"] == 0))
        check_split(d, r, null_ok=True)
    assert nnull > 0


def check_split(d, r, null_ok=False):
    """Reconstruction of A from U and V, orthonormal rows of V and
    orthogonal columns of U with the singular values as norms, per kept
    sector, descending values, the reported norm."""
    sv = d.siteV
    S = Split(d, r)
    # reconstruction: A[a, v, c] = norm_U * sum_k U[a, k, c] V[x, k, y]
    for qn, blk in _a_blocks(d):
        tup = []
        for i in S.incl:
            tup.append(S.renumber(i, qn[i])[0] if i == S.cs else qn[i])
        qv, sec = S.renumber(sv, qn[sv])
        k = r["bond_dims"][S.newid[sec]] if S.newid[sec] >= 0 else 0
        if k == 0:
            assert np.max(np.abs(blk)) < 1e-12 if blk.size else True
            continue
        U, V = S.ublock(tup), S.vblock(qv)
        assert U is not None and V is not None
        # U's legs before its cut leg (U's site and leg order)
        t0 = 1
        for i in S.incl:
            di, _ = S.dims(i, qn[i], cutdim=1)
            if i < S.cs:
                t0 *= int(np.prod(di))
            elif i == S.cs:
                t0 *= int(np.prod(di[:S.cb])) if S.cb else 1
        vd, _ = S.dims(sv, qn[sv], cutdim=1)
        v0 = int(np.prod(vd[:S.vb])) if S.vb else 1
        v2 = int(np.prod(vd[S.vb + 1:]))
        t2 = U.size // (t0 * k)
        # rows: U's non-cut legs in order; columns: V's legs
        Ur = np.moveaxis(U.reshape((t0, k, t2), order="F"), 1, 2).reshape((t0 * t2, k), order="F")
        Vc = np.moveaxis(V.reshape((v0, k, v2), order="F"), 1, 2).reshape((v0 * v2, k), order="F")
        rc = r["norm_U"] * Ur @ Vc.T
        # A's block: (sites before V, V, sites after V), column-major
        a0 = int(np.prod([np.prod(S.dims(i, qn[i], cutdim=1)[0]) for i in S.incl if i < sv] or [1]))
        a2 = rc.shape[0] // a0
        rec = np.moveaxis(rc.reshape((a0, a2, v0 * v2), order="F"), 2, 1)
        np.testing.assert_allclose(rec.flatten(order="F"), blk, rtol=0, atol=1e-12)
    # V orthonormal rows and U orthogonal columns per kept sector
    gram_v, gram_u = {}, {}
    for b in range(len(r["V_beginblock"]) - 1):
        q = r["V_qnumbers"][b]
        ids = _ids(q, [d.ss[sv][kk].nrSecs if kk != S.vb else len(S.bond_irr) for kk in range(3)])
        sec = ids[S.vb]
        k = r["bond_dims"][sec]
        V = r["V_tel"][r["V_beginblock"][b]:r["V_beginblock"][b + 1]]
        x = V.size // k
        v0 = int(np.prod([d.ss[sv][kk].dims[ids[kk]] for kk in range(S.vb)])) if S.vb else 1
        Vm = np.moveaxis(V.reshape((v0, k, x // v0), order="F"), 1, 0).reshape(k, -1)
        gram_v[sec] = gram_v.get(sec, 0) + Vm @ Vm.T
    for sec, g in gram_v.items():
        np.testing.assert_allclose(g, np.eye(g.shape[0]), rtol=0, atol=1e-11)
    nu = r["U_nrsites"]
    uq = r["U_qnumbers"].reshape(-1, nu)
    ci = S.incl.index(S.cs)
    for b in range(len(r["U_beginblock"]) - 1):
        q = int(uq[b][ci])
        ns_ = [d.ss[S.cs][kk].nrSecs if kk != S.cb else len(S.bond_irr) for kk in range(3)]
        ids = _ids(q, ns_)
        sec = ids[S.cb]
        k = r["bond_dims"][sec]
        U = r["U_tel"][r["U_beginblock"][b]:r["U_beginblock"][b + 1]]
        t0 = 1
        for j, i in enumerate(S.incl):
            di = [d.ss[i][kk].dims[x] for kk, x in enumerate(_ids(int(uq[b][j]), ns_ if i == S.cs else
                                                                    S.nsec(i)))]
            if i < S.cs:
                t0 *= int(np.prod(di))
            elif i == S.cs:
                t0 *= int(np.prod(di[:S.cb])) if S.cb else 1
        Ur = np.moveaxis(U.reshape((t0, k, U.size // (t0 * k)), order="F"), 1, 2).reshape((-1, k), order="F")
        gram_u[sec] = gram_u.get(sec, 0) + Ur.T @ Ur
    offs = np.concatenate([[0], np.cumsum(r["nsing"])])
    for new, g in gram_u.items():
        old_sec = [s_ for s_, n_ in S.newid.items() if n_ == new][0]
        sv_ = r["sing"][offs[old_sec]:offs[old_sec] + r["kept"][old_sec]] / r["norm_U"]
        np.testing.assert_allclose(g, np.diag(sv_ ** 2), rtol=0, atol=1e-12)
    for s_ in range(len(r["nsing"])):
        assert np.all(np.diff(r["sing"][offs[s_]:offs[s_ + 1]]) <= 0)
    # truncation keeps the largest values; the norm after is their weight
    if not null_ok:
        assert abs(r["norm"][0] - 1) < 1e-10


@pytest.mark.parametrize("case", [c for c in CASES if c[-2] >= 48], ids=_cid)
def test_split_truncation_keeps_largest(case):
    ctx, st = _step(case)
    d0, o, s = st.descriptors()
    for sv in splittable(st, s.nrsites):
        full = O.split_site(st.split_problem(sv, maxD=100000))
        total = int(full["nsing"].sum())
        maxD = max(1, total // 3)
        r = O.split_site(st.split_problem(sv, maxD=maxD, trunc_type="N"))
        assert int(r["kept"].sum()) == maxD
        np.testing.assert_allclose(r["sing"], full["sing"], rtol=0, atol=1e-12)
        order = np.sort(full["sing"])[::-1]
        offs = np.concatenate([[0], np.cumsum(r["nsing"])])
        kept = np.concatenate([r["sing"][offs[i]:offs[i] + r["kept"][i]] for i in range(len(r["kept"]))])
        np.testing.assert_allclose(np.sort(kept)[::-1], order[:maxD], rtol=0, atol=1e-14)
        # SU(2) sectors carry (2j+1)-fold multiplets: the norm counts the
        # reduced values as they come, U(1) values once each
        assert r["norm"][1] == pytest.approx(float(np.sum(kept ** 2)), rel=1e-12)
        assert r["norm_U"] == pytest.approx(np.sqrt(r["norm"][1]), rel=1e-12)
