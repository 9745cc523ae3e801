"""Helpers for the Hamiltonian-environment tests: the oracle's three
renormalisation calls in the form t3ns_amd.environment.Environment takes
(results carrying a RopsDesc), isometric site tensors, and exact
diagonalisation of the Hubbard model on a network's hopping graph."""
import ctypes as C
import itertools

import numpy as np

import oracle as O
from renorm_util import as_dbl, as_i64
from t3ns_amd._lib import RopsDesc


class OracleOps:
    """Oracle renormalisation output [(beginblock, elements)] per operator
    as a RopsDesc over the structure arrays of the descriptor that made it."""

    def __init__(self, blocks, is_left, P, nrhss, bbh, qnumbers, hss):
        self.blocks = blocks
        n = len(blocks)
        self._hss = np.array(hss, dtype=np.int32)
        self._bbs = [np.ascontiguousarray(b, dtype=np.int64) for b, _ in blocks]
        self._tels = [np.ascontiguousarray(t if len(t) else np.zeros(1)) for _, t in blocks]
        self._bbp = (C.POINTER(C.c_int64) * max(n, 1))(*[as_i64(b) for b in self._bbs])
        self._telp = (C.POINTER(C.c_double) * max(n, 1))(*[as_dbl(t) for t in self._tels])
        d = RopsDesc()
        d.bond = -1
        d.is_left = int(is_left)
        d.P_operator = int(P)
        d.nrhss = nrhss
        d.begin_blocks_of_hss = bbh
        d.qnumbers = qnumbers
        d.nrops = n
        d.hss_of_ops = self._hss.ctypes.data_as(C.POINTER(C.c_int))
        d.beginblock = self._bbp
        d.tel = self._telp
        self.desc = d


def oracle_append(ad, old):
    blocks = O.append_phys(ad, old)
    hss = [ad.hss_of_new[k] for k in range(len(blocks))]
    return OracleOps(blocks, ad.is_left, 1, ad.nrhss, ad.begin_blocks_of_hss, ad.qnumbers, hss)


def oracle_update_physical(pd, pops):
    blocks = O.update_physical(pd, pops)
    hss = [pops.hss_of_ops[k] for k in range(pops.nrops)]
    return OracleOps(blocks, pd.is_left, 0, pd.nrhss, pd.begin_blocks_of_hss, pd.qnumbers, hss)


def oracle_update_branching(d, ops):
    blocks = O.update_branching(d, ops)
    hss = [d.hss_of_new[k] for k in range(len(blocks))]
    return OracleOps(blocks, d.update_case == 2, 0, d.nrhss, d.begin_blocks_of_hss, d.qnumbers, hss)


ORACLE_BACKEND = dict(append=oracle_append, update_phys=oracle_update_physical,
                      update_branch=oracle_update_branching)


def isometric_site_tensor(pd, rng, keep):
    """Replaces the site tensor of update descriptor `pd` by a seeded
    isometry: for every sector of the updated bond (d states) the rows over
    all (far bond, physical) columns (c >= d of them) are orthonormal.  With
    c == d in every sector this is orthogonal_site_tensor."""
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
        if d == 0:
            continue
        sizes = [int(bb[b + 1] - bb[b]) // d for b in blks]
        c = sum(sizes)
        assert c >= d, "sector larger than its product space"
        Q, _ = np.linalg.qr(rng.standard_normal((c, d)))
        Q = Q.T  # d orthonormal rows of length c
        k = 0
        for b, m in zip(blks, sizes):
            part = Q[:, k:k + m]
            k += m
            tel[bb[b]:bb[b + 1]] = part.T.flatten(order="F") if il else part.flatten(order="F")
    keep.append(tel)
    pd.T_tel = as_dbl(tel)


def isometric_tensors(seed):
    """site_tensor hook: every physical-site update gets a seeded isometry
    (isometric_site_tensor)."""
    rng = np.random.default_rng(seed)

    def hook(site, desc, keep):
        if hasattr(desc, "update_case"):
            raise ValueError("isometric tensors are made for physical sites only")
        isometric_site_tensor(desc, rng, keep)
    return hook


def dense(matvec, n):
    """Dense matrix of a linear map on R^n, column by column."""
    H = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        H[:, j] = matvec(e)
        e[j] = 0.0
    return H


def hubbard_ground_energy(L, edges, n_up, n_down, t, U):
    """Lowest eigenvalue of the Hubbard model on the graph `edges` over L
    sites, H = -t sum_{(i,j),s} (c+_{i s} c_{j s} + h.c.) + U sum_i n_{i up}
    n_{i down}, with n_up and n_down electrons, by exact diagonalisation.
    The modes are ordered all-up then all-down (site order inside each
    spin), so a hop between sites a < b of one spin carries the
    Jordan-Wigner sign of the same-spin electrons on sites a+1..b-1."""
    def spin_part(n):
        states = [sum(1 << i for i in c) for c in itertools.combinations(range(L), n)]
        index = {s: k for k, s in enumerate(states)}
        h = np.zeros((len(states), len(states)))
        for k, s in enumerate(states):
            for i, j in edges:
                a, b = min(i, j), max(i, j)
                between = bin(s & ((1 << b) - (1 << (a + 1)))).count("1")
                for src, dst in ((a, b), (b, a)):
                    if (s >> src) & 1 and not (s >> dst) & 1:
                        h[index[s ^ (1 << src) ^ (1 << dst)], k] += -t * (-1.0) ** between
        return states, h

    su, hu = spin_part(n_up)
    sd, hd = spin_part(n_down)
    H = np.kron(hu, np.eye(len(sd))) + np.kron(np.eye(len(su)), hd)
    occ = np.array([bin(a & b).count("1") for a in su for b in sd], dtype=float)
    H += U * np.diag(occ)
    return np.linalg.eigvalsh(H)[0]


def hubbard_edges(ctx):
    """Hopping graph of NN_HUBBARD on a network: physical sites joined by a
    bond, and every pair of physical sites next to one branching site (the
    branching update couples the creation/annihilation operators of its
    legs, src/instructions_nn_hubbard.c)."""
    info = ctx.info()
    edges = set()
    for b in range(info["nr_bonds"]):
        l, r = ctx.bond_sites(b)
        if l < 0 or r < 0:
            continue
        if ctx.site_bonds(l)[1] and ctx.site_bonds(r)[1]:
            edges.add((min(l, r), max(l, r)))
    for s in range(info["sites"]):
        bonds, phys = ctx.site_bonds(s)
        if phys:
            continue
        nb = []
        for b in bonds:
            l, r = ctx.bond_sites(b)
            o = r if l == s else l
            if o < 0 or not ctx.site_bonds(o)[1]:
                raise ValueError("branching site %d has a non-physical neighbour" % s)
            nb.append(o)
        for i in range(3):
            for j in range(i + 1, 3):
                edges.add((min(nb[i], nb[j]), max(nb[i], nb[j])))
    return sorted(edges)


class OracleSolver:
    """The oracle's makesiteTensor, Heff eigenpair (dense below 800 states,
    Lanczos above) and split, in the form t3ns_amd.sweep.Sweeper takes."""

    env = ORACLE_BACKEND

    def make_site_tensor(self, desc):
        return O.make_site_tensor(desc)

    def solve(self, d, o, s, guess, tol, keep):
        import scipy.sparse.linalg as sla
        oh = O.OracleHeff(d, o, s)
        try:
            n = len(guess)
            if n <= 800:
                w, v = np.linalg.eigh(dense(oh.matvec, n))
                return w[0], v[:, 0], (0, 0)
            A = sla.LinearOperator((n, n), matvec=oh.matvec, dtype=float)
            w, v = sla.eigsh(A, k=1, which="SA", v0=guess, tol=1e-12)
            return w[0], v[:, 0], (0, 0)
        finally:
            oh.close()

    def split(self, desc):
        return O.split_site(desc)
