"""Pins of the oracle's update_rOperators_branching restatement
(src/rOperators_bUpdate.c:229-1101) that do not use its own formula.

* Contraction orders: the reference picks the cheapest of six GEMM orders
  per block (ways_to_update / fillin_cinfo, 813-1012).  Forcing each order
  for every block must give the same operators (1e-12): a wrong transpose,
  leading dimension or batch stride in any order shows.
* The unity: with identity operators on the two old bonds, the unity of
  the new bond must come out, per sector s of that bond, as X_s S X_s^T
  with X_s the rows of the site tensor in sector s over all columns of the
  two other legs and S the fermion parity of the column's leg 1 (new
  operators on leg 0), of its leg 0 (new operators on leg 1) or 1 (new
  operators on leg 2) -- the Z2 signs of prefactor_bUpdate for a unity
  (src/symmetry_z2.c:105-153).  Assembled densely from the tensor's blocks.
  U(1) cases only (the SU(2) unity carries multiplet factors)."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from conftest import netw_text
from renorm_util import set_identity
from test_qc_heff import fcidump

# (model, network, symmetries, target, D, branching site)
BCASES = [
    ("hub", "10_T3NS", ["U1", "U1"], [5, 5], 24, 10),
    ("hub", "10_T3NS", ["U1", "U1"], [5, 5], 24, 11),
    ("hub", "10_T3NS", ["U1", "SU2"], [10, 0], 24, 11),
    ("hub", "10_T3NS", ["U1", "SU2"], [10, 0], 24, 12),
    ("hub", "16_T3NS", ["U1", "SU2"], [16, 0], 48, None),
    ("qc", "ref_networks/10_T3NS", ["U1", "U1"], [7, 7], 64, 8),
    ("qc", "ref_networks/10_T3NS", ["U1", "U1", "D2h"], [7, 7, 0], 64, 2),
    ("qc", "ref_networks/10_T3NS", ["U1", "SU2", "D2h"], [14, 0, 0], 64, 8),
    ("qc", "ref_networks/10_T3NS", ["U1", "SU2"], [14, 0], 64, 11),
]


def bcase_id(c):
    return "%s-%s-%s-D%d-s%s" % (c[0], c[1], "x".join(c[2]), c[4], c[5])


def branching_sites(ctx):
    out = []
    for s in range(ctx.info()["sites"]):
        try:
            ctx.bupdate_problem(s, 2, seed=0)
            out.append(s)
        except Exception:
            pass
    return out


def context(c):
    kind, name, sy, ts, D, _ = c
    if kind == "hub":
        return T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    return T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=fcidump())


def site_of(ctx, c):
    return c[5] if c[5] is not None else branching_sites(ctx)[0]


def _close(a, b, rtol):
    assert len(a) == len(b)
    for (ab, at), (bb, bt) in zip(a, b):
        assert np.array_equal(ab, bb)
        if len(bt):
            assert np.max(np.abs(at - bt)) <= rtol * max(1.0, np.max(np.abs(bt)))


@pytest.mark.parametrize("uc", [0, 1, 2])
@pytest.mark.parametrize("case", BCASES, ids=bcase_id)
def test_bupdate_contraction_orders_agree(case, uc):
    ctx = context(case)
    d, o = ctx.bupdate_problem(site_of(ctx, case), uc, seed=5 + uc)
    ref = O.update_branching(d, o)
    assert sum(len(t) for _, t in ref) > 0
    for way in range(6):
        _close(O.update_branching(d, o, order=way), ref, 1e-12)


def _rows(d, s):
    """Site tensor rows of sector s of leg update_case over the two other
    legs, with the parity sign of each column."""
    uc = d.update_case
    ss = d.site_ss
    n0, n1 = ss[0].nrSecs, ss[1].nrSecs
    dn = ss[uc].dims[s]
    total = d.T_beginblock[d.T_nrblocks]
    tel = np.ctypeslib.as_array(d.T_tel, shape=(max(total, 1),))
    mats, signs = [], []
    for b in range(d.T_nrblocks):
        q = d.T_qnumbers[b]
        ids = (q % n0, (q // n0) % n1, q // (n0 * n1))
        if ids[uc] != s:
            continue
        dims = [ss[k].dims[ids[k]] for k in range(3)]
        blk = tel[d.T_beginblock[b]:d.T_beginblock[b + 1]].reshape(dims, order="F")
        X = np.moveaxis(blk, uc, 0).reshape(dn, -1, order="F")
        sign_leg = {0: 1, 1: 0, 2: None}[uc]
        sg = 1.0 if sign_leg is None else (-1.0) ** (ss[sign_leg].irreps[ids[sign_leg] * 5] % 2)
        mats.append(X)
        signs += [sg] * X.shape[1]
    if not mats:
        return np.zeros((dn, 0)), np.zeros(0)
    return np.hstack(mats), np.array(signs)


@pytest.mark.parametrize("uc", [0, 1, 2])
@pytest.mark.parametrize("case", [c for c in BCASES if "SU2" not in c[2]], ids=bcase_id)
def test_bupdate_unity(case, uc):
    ctx = context(case)
    d, o = ctx.bupdate_problem(site_of(ctx, case), uc, seed=7 + uc)
    keep = []
    for r in range(2):
        set_identity(o[r], d.site_ss[[1 if uc == 0 else 0, 1 if uc == 2 else 2][r]].nrSecs, keep)
    out = O.update_branching(d, o)
    ns = d.site_ss[uc].nrSecs
    dims = [d.site_ss[uc].dims[i] for i in range(ns)]
    found = 0
    for k, (bb, tel) in enumerate(out):
        h = d.hss_of_new[k]
        b0 = d.begin_blocks_of_hss[h]
        ok, ndiag = True, 0
        for b in range(len(bb) - 1):
            q = d.qnumbers[b0 + b]
            bra, ket = q % ns, (q // ns) % ns
            blk = tel[bb[b]:bb[b + 1]]
            if bra != ket:
                ok = ok and (len(blk) == 0 or np.max(np.abs(blk)) < 1e-12)
                continue
            X, sg = _rows(d, bra)
            want = (X * sg) @ X.T
            if not np.any(np.abs(want) >= 1e-20):
                ok = ok and len(blk) == 0
                continue
            ndiag += 1
            if len(blk) != dims[bra] ** 2:
                ok = False
                continue
            got = blk.reshape((dims[bra], dims[bra]), order="F")
            ok = ok and np.max(np.abs(got - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
        if ok and ndiag > 0:
            found += 1
    assert found >= 1, "no operator equals the updated unity"


def _prefactor_bupdate(d, irr, uc):
    import ctypes as C
    flat = (C.c_int * 45)(*[v for bond in irr for t in bond for v in t])
    sgs = (C.c_int * 5)(*[d.sgs[i] for i in range(5)])
    return O.lib().oracle_prefactor_bUpdate(flat, uc, sgs, d.nrSyms)


def _irrep(ss, i):
    return tuple(ss.irreps[i * 5 + k] for k in range(5))


def _dense_branching(d, o):
    """update_rOperators_branching restated densely: every (ket block T,
    adjoint block A) pair of the site tensor whose legs match an old block
    of each operator contributes pref * einsum(A, O0, O1, T) to the unique
    operator of the instruction run, then the instruction sums."""
    uc = d.update_case
    x = 1 if uc == 0 else 0
    y = 1 if uc == 2 else 2
    ss = d.site_ss
    nsec = [ss[k].nrSecs for k in range(3)]
    H = d.mpo.nrSecs
    tel = np.ctypeslib.as_array(d.T_tel, shape=(max(d.T_beginblock[d.T_nrblocks], 1),))

    def tblock(b):
        q = d.T_qnumbers[b]
        ids = (q % nsec[0], (q // nsec[0]) % nsec[1], q // (nsec[0] * nsec[1]))
        dims = [ss[k].dims[ids[k]] for k in range(3)]
        return ids, tel[d.T_beginblock[b]:d.T_beginblock[b + 1]].reshape(dims, order="F")

    def opblock(r, op, bra, ket, bond):
        h = o[r].hss_of_ops[op]
        n = nsec[bond]
        q = bra + ket * n + h * n * n
        b0, b1 = o[r].begin_blocks_of_hss[h], o[r].begin_blocks_of_hss[h + 1]
        for b in range(b1 - b0):
            if o[r].qnumbers[b0 + b] == q:
                lo, hi = o[r].beginblock[op][b], o[r].beginblock[op][b + 1]
                if hi == lo:
                    return None
                arr = np.ctypeslib.as_array(o[r].tel[op], shape=(hi,))[lo:hi]
                return arr.reshape((ss[bond].dims[bra], ss[bond].dims[ket]), order="F")
        return None

    tprods = {h: {(d.tprods[2 * p], d.tprods[2 * p + 1])
                  for p in range(d.tprod_begin[h], d.tprod_begin[h + 1])} for h in range(H)}
    runs = []
    for i in range(d.nr_instr):
        ins = tuple(d.instr[3 * i + k] for k in range(3))
        key = (ins[0], ins[1], d.hss_of_new[ins[2]])
        if not runs or runs[-1][0] != key:
            runs.append((key, []))
        runs[-1][1].append(i)
    nnew = d.begin_blocks_of_hss[H]
    uniq = {}
    blocks = [tblock(b) for b in range(d.T_nrblocks)]
    by_sector = {}
    for tid, Tb in blocks:
        by_sector.setdefault(tid[uc], []).append((tid, Tb))
    letters = ["a", "b", "c"]
    for ru, ((op0, op1, hn), _) in enumerate(runs):
        h0, h1 = o[0].hss_of_ops[op0], o[1].hss_of_ops[op1]
        if (h0, h1) not in tprods[hn]:
            continue
        for nsb in range(d.begin_blocks_of_hss[hn], d.begin_blocks_of_hss[hn + 1]):
            q = d.qnumbers[nsb]
            bn, kn = q % nsec[uc], (q // nsec[uc]) % nsec[uc]
            acc = None
            for tid, Tb in by_sector.get(kn, []):
                for aid, Ab in by_sector.get(bn, []):
                    O0 = opblock(0, op0, aid[x], tid[x], x)
                    O1 = opblock(1, op1, aid[y], tid[y], y)
                    if O0 is None or O1 is None:
                        continue
                    hs = [0, 0, 0]
                    hs[x], hs[y], hs[uc] = h0, h1, hn
                    irr = [(_irrep(ss[k], aid[k]), _irrep(ss[k], tid[k]), _irrep(d.mpo, hs[k]))
                           for k in range(3)]
                    pref = _prefactor_bupdate(d, irr, uc)
                    # A[bra legs], T[ket legs]; O0 on leg x, O1 on leg y
                    la = ["A", "B", "C"]
                    la[uc] = "n"
                    lt = list(letters)
                    lt[uc] = "m"
                    expr = "%s,%s%s,%s%s,%s->nm" % ("".join(la), la[x], lt[x], la[y], lt[y], "".join(lt))
                    c = pref * np.einsum(expr, Ab, O0, O1, Tb)
                    acc = c if acc is None else acc + c
            if acc is not None:
                uniq[(ru, nsb)] = acc
    # instruction sums and zero-block removal
    nops = max(d.instr[3 * i + 2] for i in range(d.nr_instr)) + 1
    out = []
    for k in range(nops):
        h = d.hss_of_new[k]
        nb = d.begin_blocks_of_hss[h + 1] - d.begin_blocks_of_hss[h]
        bb, tels = [0], []
        for b in range(nb):
            nsb = d.begin_blocks_of_hss[h] + b
            q = d.qnumbers[nsb]
            dims = (ss[uc].dims[q % nsec[uc]], ss[uc].dims[(q // nsec[uc]) % nsec[uc]])
            v = np.zeros(dims)
            for ru, (_, members) in enumerate(runs):
                for i in members:
                    if d.instr[3 * i + 2] == k and (ru, nsb) in uniq:
                        v = v + d.pref[i] * uniq[(ru, nsb)]
            flat = v.flatten(order="F")
            if flat.size and np.any(np.abs(flat) >= 1e-20):
                tels.append(flat)
                bb.append(bb[-1] + flat.size)
            else:
                bb.append(bb[-1])
        out.append((np.array(bb, dtype=np.int64), np.concatenate(tels) if tels else np.zeros(0)))
    return out


@pytest.mark.parametrize("uc", [0, 1, 2])
@pytest.mark.parametrize("case", [BCASES[1], BCASES[2], BCASES[6]], ids=bcase_id)
def test_bupdate_dense_restatement(case, uc):
    ctx = context(case)
    d, o = ctx.bupdate_problem(site_of(ctx, case), uc, seed=11 + uc)
    _close(O.update_branching(d, o), _dense_branching(d, o), 1e-12)
