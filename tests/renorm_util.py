"""Helpers for the renormalisation tests: operator descriptors built from
numpy arrays, identity operators and dense sector views of a site tensor."""
import ctypes as C

import numpy as np

from t3ns_amd._lib import MAX_SYMMETRIES, RopsDesc


def as_i64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


def as_dbl(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_identity(old, nbra, keep):
    """Sets every operator of `old` (ordinary operators, one quantum number
    bra + ket * nbra + hss * nbra^2 per block) to the identity on its blocks
    with equal bra and ket sector and to zero elsewhere; the arrays go to
    `keep` so that they outlive the descriptor."""
    arrs = []
    for k in range(old.nrops):
        h = old.hss_of_ops[k]
        b0, b1 = old.begin_blocks_of_hss[h], old.begin_blocks_of_hss[h + 1]
        bb = old.beginblock[k]
        t = np.zeros(max(int(bb[b1 - b0]), 1))
        for b in range(b1 - b0):
            q = old.qnumbers[b0 + b]
            sz = int(bb[b + 1] - bb[b])
            if sz and q % nbra == (q // nbra) % nbra:
                d = int(round(np.sqrt(sz)))
                t[bb[b]:bb[b + 1]] = np.eye(d).flatten()
        arrs.append(t)
    ptrs = (C.POINTER(C.c_double) * old.nrops)(*[as_dbl(t) for t in arrs])
    keep += [arrs, ptrs]
    old.tel = ptrs


def physical_rops(ad, blocks, keep):
    """RopsDesc of physical operators with the structure of append
    descriptor `ad` and per-operator (beginblock, elements) `blocks`."""
    nops = len(blocks)
    d = RopsDesc()
    d.bond = ad.new_ss[1].bond
    d.is_left = ad.is_left
    d.P_operator = 1
    d.nrhss = ad.nrhss
    d.begin_blocks_of_hss = ad.begin_blocks_of_hss
    d.qnumbers = ad.qnumbers
    d.nrops = nops
    hss = np.array([ad.hss_of_new[k] for k in range(nops)], dtype=np.int32)
    bbs = [np.ascontiguousarray(b, dtype=np.int64) for b, _ in blocks]
    tels = [np.ascontiguousarray(t if len(t) else np.zeros(1)) for _, t in blocks]
    bbp = (C.POINTER(C.c_int64) * nops)(*[as_i64(b) for b in bbs])
    telp = (C.POINTER(C.c_double) * nops)(*[as_dbl(t) for t in tels])
    keep += [hss, bbs, tels, bbp, telp]
    d.hss_of_ops = hss.ctypes.data_as(C.POINTER(C.c_int))
    d.beginblock = bbp
    d.tel = telp
    return d


def unity_ops(ad, blocks):
    """Indices of the physical operators that are the identity: blocks with
    equal bra and ket site quantum numbers are identity matrices, all other
    blocks zero or absent."""
    out = []
    for k, (bb, t) in enumerate(blocks):
        h = ad.hss_of_new[k]
        b0 = ad.begin_blocks_of_hss[h]
        ok, diag = True, False
        for b in range(len(bb) - 1):
            blk = t[bb[b]:bb[b + 1]]
            if ad.qnumbers[3 * (b0 + b)] == ad.qnumbers[3 * (b0 + b) + 1]:
                diag = True
                d = int(round(np.sqrt(len(blk))))
                if not len(blk) or not np.allclose(blk, np.eye(d).flatten(), rtol=0, atol=1e-13):
                    ok = False
            elif len(blk) and np.max(np.abs(blk)) > 1e-13:
                ok = False
        if ok and diag:
            out.append(k)
    return out


def sector_rows(pd, s):
    """The site tensor of update descriptor `pd` as a dense matrix whose
    rows are the updated bond's sector `s` and whose columns run over all
    (far bond, physical) pairs, with the fermion parity (-1)^|phys| of every
    column (the Z2 irrep is symmetry 0)."""
    il = pd.is_left
    ss = pd.site_ss
    n0, n1 = ss[0].nrSecs, ss[1].nrSecs
    d = ss[2 if il else 0].dims[s]
    mats, signs = [], []
    if pd.T_nrblocks == 0:
        return np.zeros((d, 0)), np.zeros(0)
    total = pd.T_beginblock[pd.T_nrblocks]
    tel = np.ctypeslib.as_array(pd.T_tel, shape=(max(total, 1),))
    for b in range(pd.T_nrblocks):
        q = pd.T_qnumbers[b]
        i0, i1, i2 = q % n0, (q // n0) % n1, q // (n0 * n1)
        if (i2 if il else i0) != s:
            continue
        blk = tel[pd.T_beginblock[b]:pd.T_beginblock[b + 1]]
        m = len(blk) // d
        # left: (far x phys) rows by the bond's columns; right: the bond's
        # rows by (phys x far) columns; both column-major
        mats.append(blk.reshape((m, d), order="F").T if il else blk.reshape((d, m), order="F"))
        signs += [(-1.0) ** (ss[1].irreps[i1 * MAX_SYMMETRIES] % 2)] * m
    if not mats:
        return np.zeros((d, 0)), np.zeros(0)
    return np.hstack(mats), np.array(signs)


def orthogonal_site_tensor(pd, rng, keep):
    """Replaces the site tensor of update descriptor `pd` by one whose rows
    of every sector of the updated bond, over all (far bond, physical)
    columns, form an orthogonal matrix (the sector dims must equal the
    column counts, as a decomposition without truncation leaves them)."""
    il = pd.is_left
    ss = pd.site_ss
    n0, n1 = ss[0].nrSecs, ss[1].nrSecs
    nb = pd.T_nrblocks
    qn = np.array([pd.T_qnumbers[i] for i in range(nb)], dtype=np.int64)
    bb = np.array([pd.T_beginblock[i] for i in range(nb + 1)], dtype=np.int64)
    sec = qn // (n0 * n1) if il else qn % n0
    tel = np.zeros(max(int(bb[-1]), 1))
    for s in np.unique(sec):
        blks = np.nonzero(sec == s)[0]
        d = ss[2 if il else 0].dims[int(s)]
        sizes = [int(bb[b + 1] - bb[b]) // d for b in blks]
        assert sum(sizes) == d, "sector is truncated"
        Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
        c = 0
        for b, m in zip(blks, sizes):
            part = Q[:, c:c + m]  # the bond's rows by this block's columns
            c += m
            tel[bb[b]:bb[b + 1]] = part.T.flatten(order="F") if il else part.flatten(order="F")
    keep.append(tel)
    pd.T_tel = as_dbl(tel)


def symsecs_dims(ss):
    return [ss.dims[i] for i in range(ss.nrSecs)]


def isometric_blocks(qn, bb, ss, leg, rng):
    """Elements of a one-site tensor (blocks `qn`, offsets `bb`, sectors
    `ss` of its three bonds) whose rows over leg `leg` (0 or 2) are
    orthonormal within every sector of that leg, over all blocks of the
    sector (columns: the other two legs).  Returns (elements, sectors where
    that was possible: column count >= sector dim)."""
    n0, n1 = ss[0].nrSecs, ss[1].nrSecs
    qn = np.asarray(qn, dtype=np.int64)
    bb = np.asarray(bb, dtype=np.int64)
    sec = qn % n0 if leg == 0 else qn // (n0 * n1)
    tel = rng.standard_normal(max(int(bb[-1]), 1))
    good = set()
    for s in np.unique(sec):
        blks = np.nonzero(sec == s)[0]
        d = ss[leg].dims[int(s)]
        if d == 0:
            continue
        sizes = [int(bb[b + 1] - bb[b]) // d for b in blks]
        R = sum(sizes)
        if R < d:
            continue
        good.add(int(s))
        Q, _ = np.linalg.qr(rng.standard_normal((R, d)))
        X = Q.T  # d x R, orthonormal rows
        c = 0
        for b, m in zip(blks, sizes):
            part = X[:, c:c + m]
            c += m
            # leg 0: block is (d x m) column-major; leg 2: (m x d)
            tel[bb[b]:bb[b + 1]] = part.flatten(order="F") if leg == 0 else part.T.flatten(order="F")
    return tel, good
