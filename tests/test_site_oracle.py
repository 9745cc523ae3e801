"""Pins of the oracle's makesiteTensor restatement (src/siteTensor_init.c:
609-800) that do not use its GEMM sequence.

* Dense contraction: every multisite block recomputed with numpy.einsum over
  explicitly labelled legs of its one-site blocks and laid out in site
  order (sites as the multisite lists them, legs in order, internal legs
  dropped), column-major.  Pins the reference's layout, including the
  transposed and batched products it uses when a physical site hangs on
  the branching site's middle leg.
* Isometries: with every outer site's rows over its internal bond
  orthonormal within each sector, each inner-site block keeps its norm
  summed over the multisite blocks built from it.  Pins the block lookup
  (translation of the internal sectors, block search) independently of the
  einsum check, which shares that lookup."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from conftest import netw_text
from renorm_util import as_dbl, isometric_blocks
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


def _ids(qn, ss):
    n0, n1 = ss[0].nrSecs, ss[1].nrSecs
    return qn % n0, (qn // n0) % n1, qn // (n0 * n1)


def _irrep(ss, i):
    return tuple(ss.irreps[i * 5 + k] for k in range(5))


def _site_block(md, s, nqn):
    """(one-site block index, its dims) of site s for multisite qn nqn."""
    nss, oss = md.new_ss[s], md.old_ss[s]
    nid = _ids(nqn, nss)
    oid = []
    for k in range(3):
        want = _irrep(nss[k], nid[k])
        hit = [i for i in range(oss[k].nrSecs) if _irrep(oss[k], i) == want]
        if not hit:
            return None
        oid.append(hit[0])
    oqn = oid[0] + oid[1] * oss[0].nrSecs + oid[2] * oss[0].nrSecs * oss[1].nrSecs
    qns = [md.site_qnumbers[s][b] for b in range(md.site_nrblocks[s])]
    if oqn not in qns:
        return None
    b = qns.index(oqn)
    if md.site_beginblock[s][b + 1] == md.site_beginblock[s][b]:
        return None
    return b, [oss[k].dims[oid[k]] for k in range(3)]


def _site_array(md, s, b, dims):
    lo, hi = md.site_beginblock[s][b], md.site_beginblock[s][b + 1]
    return np.ctypeslib.as_array(md.site_tel[s], shape=(hi,))[lo:hi].reshape(dims, order="F")


def _dense_block(md, sb):
    ns = md.nrsites
    bonds = [[md.new_ss[s][k].bond for k in range(3)] for s in range(ns)]
    shared = {b for s in range(ns) for b in bonds[s]
              if sum(b in bonds[t] for t in range(ns)) == 2}
    labels, ops, out = {}, [], []
    letters = iter("abcdefghijklmnopqrstuvwxyz")
    for s in range(ns):
        hit = _site_block(md, s, md.qnumbers[sb * ns + s])
        if hit is None:
            return None
        b, dims = hit
        subs = ""
        for k in range(3):
            key = ("int", bonds[s][k]) if bonds[s][k] in shared else ("ext", s, k)
            if key not in labels:
                labels[key] = next(letters)
            subs += labels[key]
            if key[0] == "ext":
                out.append(labels[key])
        ops.append((subs, _site_array(md, s, b, dims)))
    expr = ",".join(o[0] for o in ops) + "->" + "".join(out)
    return np.einsum(expr, *[o[1] for o in ops]).flatten(order="F")


@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_makesite_dense_einsum(case):
    ctx, st = _step(case)
    md = st.makesite_problem(seed=7)
    got = O.make_site_tensor(md)
    nonzero = 0
    for sb in range(md.nrblocks):
        lo, hi = md.beginblock[sb], md.beginblock[sb + 1]
        want = _dense_block(md, sb)
        if want is None:
            assert not np.any(got[lo:hi])
            continue
        assert len(want) == hi - lo
        np.testing.assert_allclose(got[lo:hi], want, rtol=1e-12, atol=1e-12)
        nonzero += 1
    assert nonzero > 0


@pytest.mark.parametrize("case", CASES, ids=_cid)
def test_makesite_isometries_keep_inner_norms(case):
    ctx, st = _step(case)
    md = st.makesite_problem(seed=8)
    I = md.innersite
    ns = md.nrsites
    rng = np.random.default_rng(3)
    keep, good, conn = [], {}, {}
    for s in range(ns):
        if s == I:
            continue
        for j in range(3):
            for k in (0, 2):
                if md.new_ss[s][k].bond == md.new_ss[I][j].bond:
                    conn[s] = (j, k)
        j, k = conn[s]
        n = md.site_nrblocks[s]
        qn = [md.site_qnumbers[s][b] for b in range(n)]
        bb = [md.site_beginblock[s][b] for b in range(n + 1)]
        tel, good[s] = isometric_blocks(qn, bb, md.old_ss[s], k, rng)
        keep.append(tel)
        md.site_tel[s] = as_dbl(tel)
    got = O.make_site_tensor(md)
    per_inner = {}
    for sb in range(md.nrblocks):
        ok = True
        for s in range(ns):
            if s == I:
                continue
            j, k = conn[s]
            hit = _site_block(md, s, md.qnumbers[sb * ns + s])
            if hit is None:
                ok = False
                break
        inner = _site_block(md, I, md.qnumbers[sb * ns + I])
        if inner is None or not ok:
            continue
        lo, hi = md.beginblock[sb], md.beginblock[sb + 1]
        per_inner[inner[0]] = per_inner.get(inner[0], 0.0) + float(np.sum(got[lo:hi] ** 2))
    nchecked = 0
    n0, n1 = md.old_ss[I][0].nrSecs, md.old_ss[I][1].nrSecs
    for b, val in per_inner.items():
        qn = md.site_qnumbers[I][b]
        ids = (qn % n0, (qn // n0) % n1, qn // (n0 * n1))
        if not all(ids[conn[s][0]] in good[s] for s in conn):
            continue
        lo, hi = md.site_beginblock[I][b], md.site_beginblock[I][b + 1]
        blk = np.ctypeslib.as_array(md.site_tel[I], shape=(hi,))[lo:hi]
        assert val == pytest.approx(float(np.sum(blk ** 2)), rel=1e-12)
        nchecked += 1
    assert nchecked > 0
