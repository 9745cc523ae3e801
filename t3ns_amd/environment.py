"""Renormalised operators of a network built from the Hamiltonian itself.

init_rops (src/optimize_network.c:428-468) makes the operators of every bond
from the vacuum at the borders inwards: a bond whose site on the covered side
is physical gets rOperators_append_phys of that site to the operators of the
site's far bond, then update_rOperators_physical with the site's tensor; a
bond whose site is a branching site gets update_rOperators_branching of the
operators on the site's two other bonds.  The operators of an optimisation
step (src/optimize_network.c:164-176) are those of its external bonds, with
the physical sites inside the step appended where the step's leg is a
physical-operator leg.

Everything runs through the library's device entry points; the operators are
held on the host between calls as the reference holds them.  Any site
tensors serve: the environment of a bond is T^dagger O T for whatever tensors
the updates used, so the effective Hamiltonian of a step is symmetric, and
with isometric tensors and no truncation it is the Hamiltonian itself on the
step's space.
"""
import numpy as np

from .heff import append_phys, update_branching, update_physical


def normalised_rows(site, desc, keep):
    """site_tensor hook: scales the seeded tensor of an update descriptor so
    that every state of the updated bond (leg 2 of a left physical update,
    leg 0 of a right one, leg `update_case` of a branching update) has unit
    norm over the tensor's other two legs.  Rows of N(0,1) entries in high
    dimension are nearly orthogonal, so the tensor is close to an isometry
    and the operators keep the Hamiltonian's scale through any number of
    updates (unscaled N(0,1) tensors multiply it by about the row length at
    every update)."""
    leg = desc.update_case if hasattr(desc, "update_case") else (2 if desc.is_left else 0)
    ss = desc.site_ss
    n0, n1 = ss[0].nrSecs, ss[1].nrSecs
    nb = desc.T_nrblocks
    if nb == 0:
        return
    qn = np.ctypeslib.as_array(desc.T_qnumbers, shape=(nb,)).astype(np.int64)
    bb = np.ctypeslib.as_array(desc.T_beginblock, shape=(nb + 1,)).astype(np.int64)
    tel = np.ctypeslib.as_array(desc.T_tel, shape=(max(int(bb[-1]), 1),))
    ids = np.stack([qn % n0, (qn // n0) % n1, qn // (n0 * n1)], axis=1)
    dims = [np.array([ss[k].dims[i] for i in range(ss[k].nrSecs)], dtype=np.int64) for k in range(3)]
    other = tuple(a for a in range(3) if a != leg)
    sq = {}
    views = []
    for b in range(nb):
        n = int(bb[b + 1] - bb[b])
        if n == 0:
            continue
        shape = tuple(int(dims[k][ids[b, k]]) for k in range(3))
        v = tel[bb[b]:bb[b + 1]].reshape(shape, order="F")
        sec = int(ids[b, leg])
        part = (v * v).sum(axis=other)
        sq[sec] = sq[sec] + part if sec in sq else part
        views.append((v, sec))
    for v, sec in views:
        scale = 1.0 / np.sqrt(np.maximum(sq[sec], 1e-300))
        shp = [1, 1, 1]
        shp[leg] = -1
        v *= scale.reshape(shp)


class Environment:
    """Operators of bonds (both sides) of a context's network, built on demand.

    `site_tensor(site, desc, keep)`, if given, may replace the seeded site
    tensor of an update descriptor (desc.T_tel) before the update runs; the
    arrays it makes go to `keep`.  `append`, `update_phys` and
    `update_branch` are the three renormalisation calls (the device ones by
    default); each returns an object with a `desc` (RopsDesc) attribute."""

    def __init__(self, ctx, seed=0, site_tensor=None, append=append_phys,
                 update_phys=update_physical, update_branch=update_branching):
        self.ctx = ctx
        self.seed = int(seed)
        self.site_tensor = site_tensor
        self._append, self._uphys, self._ubranch = append, update_phys, update_branch
        self._ops = {}    # key -> (operators, arrays they were made from)
        self._cover = {}  # (bond, side) -> sites on that side
        self._count = 0

    def _next_seed(self):
        self._count += 1
        return self.seed * 1000003 + self._count

    def _outer_site(self, bond, is_left):
        left, right = self.ctx.bond_sites(bond)
        return left if is_left else right

    def is_border(self, bond, is_left):
        """True when the covered side of `bond` is empty (vacuum operators)."""
        return self._outer_site(bond, is_left) == -1

    def physical(self, bond, is_left):
        """Physical operators on `bond`: the site on side `is_left` appended to
        the operators of its far bond (rOperators_append_phys).  Not cached:
        they live on the bond's current sectors, which differ between a step
        where the bond is internal (the product sectors of the step's
        tensor) and one where it is not."""
        L = int(bool(is_left))
        site = self._outer_site(bond, L)
        bonds, phys = self.ctx.site_bonds(site)
        if not phys:
            raise ValueError("site %d next to bond %d is not physical" % (site, bond))
        ad, old = self.ctx.append_problem_at(bond, L, self._next_seed())
        far = bonds[0] if L else bonds[2]
        if not self.is_border(far, L):
            old = self.operators(far, L).desc
        r = self._append(ad, old)
        r.desc.bond = bond
        r.made_from = ad
        return r

    def operators(self, bond, is_left):
        """Renormalised operators on `bond` covering its side `is_left`."""
        key = ("R", int(bond), int(bool(is_left)))
        if key in self._ops:
            return self._ops[key][0]
        L = int(bool(is_left))
        site = self._outer_site(bond, L)
        if site == -1:
            raise ValueError("bond %d has a border on that side (vacuum operators)" % bond)
        bonds, phys = self.ctx.site_bonds(site)
        keep = []
        if phys:
            P = self.physical(bond, L)
            pd = self.ctx.pupdate_problem_at(bond, L, self._next_seed())
            if self.site_tensor is not None:
                self.site_tensor(site, pd, keep)
            r = self._uphys(pd, P.desc)
            keep.append(pd)
        else:
            case = bonds.index(bond)
            d, o = self.ctx.bupdate_problem(site, case, self._next_seed())
            # roles of the two old operators (src/rOperators_bUpdate.c:292-313)
            obond = (bonds[1 if case == 0 else 0], bonds[1 if case == 2 else 2])
            oleft = (1, int(case == 2))
            for k in range(2):
                if not self.is_border(obond[k], oleft[k]):
                    o[k] = self.operators(obond[k], oleft[k]).desc
            if self.site_tensor is not None:
                self.site_tensor(site, d, keep)
            r = self._ubranch(d, o)
            keep.append((d, o))
        r.desc.bond = bond
        self._ops[key] = (r, keep)
        return r

    def covered(self, bond, is_left):
        """The sites on side `is_left` of `bond`."""
        key = (int(bond), int(bool(is_left)))
        if key not in self._cover:
            nr_bonds = self.ctx.info()["nr_bonds"]
            start = self._outer_site(bond, is_left)
            seen, stack = set(), [start] if start != -1 else []
            while stack:
                s = stack.pop()
                if s in seen:
                    continue
                seen.add(s)
                for b in self.ctx.site_bonds(s)[0]:
                    if b == bond or b < 0 or b >= nr_bonds:
                        continue
                    for o in self.ctx.bond_sites(b):
                        if o != -1 and o not in seen:
                            stack.append(o)
            self._cover[key] = frozenset(seen)
        return self._cover[key]

    def invalidate(self, sites):
        """Drops every operator set that covers one of `sites` (their
        tensors changed); they are rebuilt on the next request."""
        sites = set(sites)
        for key in list(self._ops):
            if self.covered(key[1], key[2]) & sites:
                del self._ops[key]

    def step_operators(self, step):
        """(desc, ops[3], site) of `step` with its operators replaced by the
        Hamiltonian's: physical-operator legs get the append of their step
        site to the operators of the external bond, other legs those
        operators themselves, border legs stay the vacuum."""
        d, o, s = step.descriptors()
        self._step_ops = []
        for i in range(3):
            if o[i].nrhss <= 0:
                continue
            L = o[i].is_left
            if o[i].P_operator:
                r = self.physical(o[i].bond, L)
                self._step_ops.append(r)  # alive until the next step's request
                o[i] = r.desc
            elif not self.is_border(o[i].bond, L):
                o[i] = self.operators(o[i].bond, L).desc
        return d, o, s
