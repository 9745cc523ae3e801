"""Block helpers for the qr_step tests: sector matrices of a one-site
tensor over one of its bonds, and the two-site contraction over a bond."""
import numpy as np


def ids3(q, ss):
    n0, n1 = ss[0].nrSecs, ss[1].nrSecs
    return q % n0, (q // n0) % n1, q // (n0 * n1)


def shared_bond(d):
    for i in range(3):
        for j in range(3):
            if d.ssA[i].bond == d.ssC[j].bond:
                return i, j


def blocks(nrblocks, qn, bb, tel, ss, dims_override=None):
    """[(qn, ids, dims, array)] of a one-site tensor; dims_override maps a
    bond position to per-sector dims replacing the symsecs' (the new bond
    dims after the QR)."""
    out = []
    for b in range(nrblocks):
        q = int(qn[b])
        ids = ids3(q, ss)
        dims = [ss[k].dims[ids[k]] for k in range(3)]
        if dims_override:
            for k, dv in dims_override.items():
                dims[k] = int(dv[ids[k]])
        lo, hi = int(bb[b]), int(bb[b + 1])
        if hi == lo:
            continue
        out.append((q, ids, dims, np.asarray(tel[lo:hi]).reshape(dims, order="F")))
    return out


def sector_columns(blks, leg):
    """Per sector of `leg`: the blocks' elements as rows (other legs) by
    columns (the leg), stacked in block order."""
    out = {}
    for q, ids, dims, a in blks:
        m = np.moveaxis(a, leg, 2).reshape((-1, dims[leg]), order="F")
        out.setdefault(ids[leg], []).append(m)
    return {s: np.vstack(v) for s, v in out.items()}


def contract_pair(ablks, cblks, la, lc):
    """sum over the shared bond of A x C, keyed by the two blocks' other
    legs' sector indices."""
    out = {}
    for qa, ia, da, a in ablks:
        for qc, ic, dc, c in cblks:
            if ia[la] != ic[lc]:
                continue
            t = np.tensordot(a, c, axes=([la], [lc]))
            key = (tuple(x for k, x in enumerate(ia) if k != la), tuple(x for k, x in enumerate(ic) if k != lc))
            out[key] = out.get(key, 0) + t
    return out
