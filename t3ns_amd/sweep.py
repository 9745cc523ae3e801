"""Sweeps: DMRG on chain networks and T3NS sweeps on trees.

The optimisation loop of src/optimize_network.c:488-570 (execute_optScheme,
optimize_network: one step per site pair of the sweep order of
src/network.c:529-727) for the two-site DMRG case, on the library's entry
points: per step, makesiteTensor of the pair from its one-site tensors
(src/siteTensor_init.c:785-800), the step's operators
(src/optimize_network.c:164-176, through Environment), the lowest
eigenpair of its effective Hamiltonian (Davidson, FP64 or FP32 with FP64
refinement), decompose_siteTensor with truncation by repeated splits
(src/siteTensor_decompose.c:1403-1444, 1533-1585) that leave a site shared
with the next step as the orthogonality centre, each cut bond's new
sectors written into the context (the bookkeeper update of
adapt_UV_tensors_and_kick_empties), and the operators covering the step's
sites dropped so that the next step renormalises them afresh
(postprocess_rOperators, src/optimize_network.c:236-311).

The starting state is seeded and in canonical form toward the first step:
every other site tensor is an isometry onto its bond toward the first pair
(the reference starts from a random state and a renormalisation pass,
src/optimize_network.c:574-640).
"""

import time
from collections import Counter, namedtuple

import numpy as np

from . import _lib
from .environment import Environment
from .heff import FP64, Heff, Step, davidson_refined, make_site_tensor, split_site


# One regime of an optimisation scheme, fields in the order of struct regime
# (include/optScheme.h:28-45): the truncation (minD, maxD, truncerr of
# struct SvalSelect), the Davidson tolerance and iteration limit, the sweep
# limit and the energy convergence between sweeps.  (The reference's
# sitesize and noise fields are the Sweeper's site_size and zero here.)
Regime = namedtuple("Regime", ["min_dim", "max_dim", "truncerr", "davidson_tol",
                               "davidson_max_its", "max_sweeps", "e_conv"])


class SiteData:
    """A one-site tensor on the host: qnumbers, beginblock, elements."""

    def __init__(self, qn, bb, tel):
        self.qn = np.ascontiguousarray(qn, dtype=np.int64)
        self.bb = np.ascontiguousarray(bb, dtype=np.int64)
        self.tel = np.ascontiguousarray(tel if len(tel) else np.zeros(1), dtype=np.float64)

    @property
    def nrblocks(self):
        return len(self.qn)

    def pointers(self):
        return (self.qn.ctypes.data_as(_lib.c_i64_p), self.bb.ctypes.data_as(_lib.c_i64_p),
                self.tel.ctypes.data_as(_lib.c_dbl_p))


def _site_dims(ctx, site):
    bonds, _ = ctx.site_bonds(site)
    return [np.array(ctx.bond_dims(b), dtype=np.int64) for b in bonds]


def isometrise(ctx, site, data, leg, rng=None):
    """Makes `data` an isometry onto leg `leg`: for every sector of that
    leg (d states) the rows over all the tensor's other index pairs (c >= d
    of them) are orthonormal.  With `rng` they are drawn from it; without,
    they are the Q factor of the QR decomposition of the tensor's own rows
    in that sector (the canonicalisation of a given state, as the
    reference's qr of its random start, src/optimize_network.c:493-515)."""
    dims = _site_dims(ctx, site)
    n0, n1 = len(dims[0]), len(dims[1])
    ids = np.stack([data.qn % n0, (data.qn // n0) % n1, data.qn // (n0 * n1)], axis=1)
    views = {}
    for b in range(data.nrblocks):
        if data.bb[b + 1] == data.bb[b]:
            continue
        shape = tuple(int(dims[k][ids[b, k]]) for k in range(3))
        v = data.tel[data.bb[b]:data.bb[b + 1]].reshape(shape, order="F")
        views.setdefault(int(ids[b, leg]), []).append(np.moveaxis(v, leg, 0))
    for sec, vs in views.items():
        d = vs[0].shape[0]
        sizes = [v[0].size for v in vs]
        c = sum(sizes)
        if c < d:
            raise ValueError("sector %d of site %d is larger than its product space" % (sec, site))
        if rng is not None:
            M = rng.standard_normal((c, d))
        else:
            M = np.concatenate([v.reshape(d, -1) for v in vs], axis=1).T
        Q, _ = np.linalg.qr(M)
        rows = Q.T
        k = 0
        for v, m in zip(vs, sizes):
            v[...] = rows[:, k:k + m].reshape(v.shape)
            k += m


class DeviceSolver:
    """The device entry points for the steps: makesiteTensor, Heff +
    Davidson (precision "f64" or "mixed"), split."""

    env = {}

    def __init__(self, precision="f64", max_vecs=30, keep_deflate=2, max_its=300,
                 mixed_min_size=200000, autotune_min_size=None):
        """mixed_min_size: below this many elements a step's Heff is
        launch-bound, FP32 matvecs save nothing and the refinement adds
        matvecs, so "mixed" solves such steps in FP64.  autotune_min_size:
        steps of at least this many elements have their work lists
        autotuned (Heff.autotune; shapes seen before reuse their setting);
        None (default) keeps the default setting: a step's shape rarely
        repeats exactly within a sweep, and six timed settings cost about
        as much as a step's Davidson at QC-24 D = 128."""
        if precision not in ("f64", "mixed"):
            raise ValueError("precision is f64 or mixed")
        self.precision = precision
        self.max_vecs, self.keep, self.max_its = max_vecs, keep_deflate, max_its
        self.mixed_min_size = int(mixed_min_size)
        self.autotune_min_size = autotune_min_size
        # steps solved with FP32 matvecs (plus FP64 refinement) and in FP64
        # alone, and the Davidson iterations of each kind
        self.counts = {"steps_f32": 0, "steps_f64": 0, "its_f32": 0, "its_f64": 0}
        # seconds in the Heff set-up (plan, upload: the reference's "Heff:
        # prepare data") and in Davidson (diagonal and matvecs)
        self.timers = Counter()

    def make_site_tensor(self, desc):
        return make_site_tensor(desc)[0]

    def solve(self, d, o, s, guess, tol, keep):
        t0 = time.perf_counter()
        h64 = Heff(d, o, s, FP64, keepalive=keep)
        self.timers["Heff: prepare data"] += time.perf_counter() - t0
        if self.autotune_min_size is not None and h64.size >= self.autotune_min_size:
            t0 = time.perf_counter()
            h64.autotune()
            self.timers["Heff: autotune"] += time.perf_counter() - t0
        t0 = time.perf_counter()
        try:
            if self.precision == "f64" or h64.size < self.mixed_min_size:
                e, x, its, rc = h64.davidson(guess, self.max_vecs, self.keep, tol, self.max_its)
                self.counts["steps_f64"] += 1
                self.counts["its_f64"] += its
                self.timers["Heff: davidson"] += time.perf_counter() - t0
                return e, x, (0, its)
            # the FP32 operator on the FP64 object's plan and lists, its
            # arenas converted on the device (no second host set-up)
            t1 = time.perf_counter()
            h32 = h64.derive_fp32()
            self.timers["Heff: prepare data"] += time.perf_counter() - t1
            try:
                t2 = time.perf_counter()
                e, x, i1, i2, rc = davidson_refined(h32, h64, guess, self.max_vecs, self.keep, tol,
                                                    tol, self.max_its)
                self.counts["steps_f32"] += 1
                self.counts["its_f32"] += i1
                self.counts["its_f64"] += i2
                self.timers["Heff: davidson"] += time.perf_counter() - t2
                return e, x, (i1, i2)
            finally:
                h32.close()
        finally:
            h64.close()

    def split(self, desc):
        return split_site(desc)[0]


class Sweeper:
    """Sweeps over the sweep order of a network for steps of up to
    `site_size` sites (src/network.c:529-727): two-site DMRG on chains; on
    trees, two-site steps and, with site_size 3 or 4, the {physical,
    branching, physical} objects where the sweep passes a branching site.

    `solver` provides make_site_tensor(desc), solve(d, o, s, guess, tol,
    keep) -> (energy, vector, iterations) and split(desc) -> split dict,
    and `solver.env` the renormalisation calls for the Environment (all
    device calls by default)."""

    def __init__(self, ctx, max_dim, seed=0, solver=None, tol=1e-8, truncerr=0.0, min_dim=0,
                 site_size=2, start="seeded", tensors=None):
        """start: "seeded" -- every site tensor from the context's seeded
        fill, the ones outside the first step replaced by random isometries
        toward it; "normal" -- every block entry N(0, 1) from one PCG64
        generator seeded `seed`, in site order then block order, the sites
        outside the first step then made isometries toward it by a QR of
        their own entries per sector.  tensors: {site: SiteData} to start
        from instead (a state in canonical form toward the first step, such
        as a checkpoint written after whole sweeps)."""
        if start not in ("seeded", "normal"):
            raise ValueError("start is seeded or normal")
        info = ctx.info()
        self.ctx, self.max_dim, self.tol, self.truncerr = ctx, int(max_dim), float(tol), truncerr
        self.min_dim = int(min_dim)  # struct SvalSelect: minD, maxD, truncerr
        self.solver = solver if solver is not None else DeviceSolver()
        self.seed = int(seed)
        self.steps = ctx.sweep_steps(int(site_size))
        toward = self._legs_toward(self.steps[0], info["nr_bonds"])
        rng = np.random.default_rng(seed)
        self.tensors = {}
        for s in range(info["sites"]):
            if tensors is not None:
                t = tensors[s]
                self.tensors[s] = SiteData(t.qn.copy(), t.bb.copy(), t.tel.copy())
                continue
            data = SiteData(*ctx.site_tensor(s, seed * 7919 + s + 1))
            if start == "normal":
                data.tel[:] = rng.standard_normal(len(data.tel))
                if s in toward:
                    isometrise(ctx, s, data, toward[s])
            elif s in toward:
                isometrise(ctx, s, data, toward[s], rng)
            self.tensors[s] = data
        # every element the steps use comes from here: the context's problem
        # builders give structure only from now on
        ctx.fill_seeded(False)
        self.env = Environment(ctx, seed, site_tensor=self._site_tensor, **self.solver.env)
        self.energies = []
        # seconds per phase, named as the reference's timers
        # (src/optimize_network.c:43-58); the solve covers Heff set-up and
        # Davidson (the solver may split it further)
        self.timers = Counter()

    def save(self, path):
        """Checkpoint of the current state (network, bond sectors,
        Hamiltonian, site tensors) in the reference's HDF5 layout
        (t3ns_amd.checkpoint); written between sweeps it restores to the
        same calculation."""
        from .checkpoint import write_checkpoint
        write_checkpoint(path, self.ctx, self.tensors)

    @classmethod
    def restore(cls, path, max_dim, **kw):
        """A Sweeper continuing from a checkpoint written by save() (or by
        the reference's write_to_disk after whole sweeps)."""
        from .checkpoint import read_checkpoint
        ctx, tensors, _ = read_checkpoint(path)
        return cls(ctx, max_dim, tensors=tensors, **kw)

    def _legs_toward(self, sites, nr_bonds):
        """For every site outside `sites`: the leg of its bond on the path to
        them (breadth-first from the step outwards)."""
        toward, frontier, seen = {}, list(sites), set(sites)
        while frontier:
            nxt = []
            for u in frontier:
                for b in self.ctx.site_bonds(u)[0]:
                    if b < 0 or b >= nr_bonds:
                        continue
                    for t in self.ctx.bond_sites(b):
                        if t == -1 or t in seen:
                            continue
                        seen.add(t)
                        toward[t] = self.ctx.site_bonds(t)[0].index(b)
                        nxt.append(t)
            frontier = nxt
        return toward

    def _site_tensor(self, site, desc, keep):
        data = self.tensors[site]
        qn, bb, tel = data.pointers()
        desc.T_nrblocks = data.nrblocks
        desc.T_qnumbers, desc.T_beginblock, desc.T_tel = qn, bb, tel
        keep.append(data)

    def _centre_site(self, k):
        """The site of step k that keeps the orthogonality centre: one the
        next step also holds (or, where the sweep turns on the same sites,
        the step after)."""
        sites = self.steps[k]
        n = len(self.steps)
        for j in (1, 2):
            nxt = self.steps[(k + j) % n]
            common = [t for t in sites if t in nxt]
            if 0 < len(common) < len(sites):
                return common[0]
        return sites[0]

    def _common_bond(self, a, b):
        return (set(self.ctx.site_bonds(a)[0]) & set(self.ctx.site_bonds(b)[0])).pop()

    def _decompose(self, sd, order, centre):
        """decompose_siteTensor by repeated splits (the HOSVD of
        src/siteTensor_decompose.c:1533-1585): a site of the remaining
        object that is a leaf of it and not the centre is split off (an
        isometry onto the cut bond), the cut bond's kept sectors go into the
        context, and the rest (singular values in) is split again until the
        centre alone remains.  Returns the cut bonds' dimensions."""
        remaining = list(order)
        dims = []
        keep = []
        while len(remaining) > 1:
            leaf = next(t for t in remaining if t != centre and sum(
                1 for u in remaining if u != t and self._adjacent(t, u)) == 1)
            nb = next(u for u in remaining if u != leaf and self._adjacent(leaf, u))
            sd.siteV = remaining.index(leaf)
            res = self.solver.split(sd)
            cut = self._common_bond(leaf, nb)
            self.ctx.set_bond(cut, res["bond_irreps"], res["bond_dims"])
            dims.append(int(np.sum(res["bond_dims"])))
            self.tensors[leaf] = SiteData(res["V_qnumbers"], res["V_beginblock"], res["V_tel"])
            idx = remaining.index(leaf)
            remaining.pop(idx)
            if len(remaining) == 1:
                self.tensors[remaining[0]] = SiteData(res["U_qnumbers"], res["U_beginblock"],
                                                      res["U_tel"])
                break
            # the split of what is left: the remaining sites' sectors, the cut
            # bond's replaced by its kept ones
            irr = np.ascontiguousarray(res["bond_irreps"], dtype=np.int32).reshape(-1)
            dm = np.ascontiguousarray(res["bond_dims"], dtype=np.int32)
            U = (np.ascontiguousarray(res["U_qnumbers"]), np.ascontiguousarray(res["U_beginblock"]),
                 np.ascontiguousarray(res["U_tel"] if len(res["U_tel"]) else np.zeros(1)))
            keep += [irr, dm, U]
            nsd = _lib.SplitDesc()
            nsd.nrSyms, nsd.sgs = sd.nrSyms, sd.sgs
            nsd.nrsites = len(remaining)
            rows = [i for i in range(len(remaining) + 1) if i != idx]
            for j, i in enumerate(rows):
                for leg in range(3):
                    ssd = sd.ss[i][leg]
                    if ssd.bond == cut:
                        ssd = _lib.SymsecsDesc(cut, len(dm), irr.ctypes.data_as(_lib.c_int_p),
                                               dm.ctypes.data_as(_lib.c_int_p))
                    nsd.ss[j][leg] = ssd
            nsd.nrblocks = len(U[0]) // len(remaining)
            nsd.qnumbers = U[0].ctypes.data_as(_lib.c_i64_p)
            nsd.beginblock = U[1].ctypes.data_as(_lib.c_i64_p)
            nsd.tel = U[2].ctypes.data_as(_lib.c_dbl_p)
            nsd.minD, nsd.maxD, nsd.truncerr, nsd.truncType = sd.minD, sd.maxD, sd.truncerr, sd.truncType
            sd = nsd
        return dims

    def _adjacent(self, a, b):
        return bool(set(self.ctx.site_bonds(a)[0]) & set(self.ctx.site_bonds(b)[0]))

    def step(self, k):
        sites = self.steps[k]
        tm = self.timers
        t0 = time.perf_counter()
        st = Step(self.ctx, sites, seed=self.seed)
        try:
            d, o, s = st.descriptors()
            order = [s.sites[i] for i in range(s.nrsites)]
            mk = st.makesite_problem(self.seed)
            for i, site in enumerate(order):
                data = self.tensors[site]
                qn, bb, tel = data.pointers()
                mk.site_nrblocks[i] = data.nrblocks
                mk.site_qnumbers[i], mk.site_beginblock[i], mk.site_tel[i] = qn, bb, tel
            x0 = np.array(self.solver.make_site_tensor(mk), dtype=np.float64)
            x0 /= np.linalg.norm(x0)
            t1 = time.perf_counter()
            tm["siteTensor: make multisite tensor"] += t1 - t0
            d, o, s = self.env.step_operators(st)
            t2 = time.perf_counter()
            tm["rOperators: append and update"] += t2 - t1
            e, x, its = self.solver.solve(d, o, s, x0, self.tol, (st, self.env))
            t3 = time.perf_counter()
            tm["Heff: solve"] += t3 - t2
            centre = self._centre_site(k)
            sd = st.split_problem(0, self.max_dim, self.min_dim, self.truncerr)
            xn = np.ascontiguousarray(x / np.linalg.norm(x))
            sd.tel = xn.ctypes.data_as(_lib.c_dbl_p)
            dims = self._decompose(sd, order, centre)
            self.env.invalidate(sites)
            tm["siteTensor: decompose"] += time.perf_counter() - t3
            return {"sites": list(sites), "energy": e, "its": its, "bond_dim": max(dims)}
        finally:
            st.close()

    def sweep(self):
        """One sweep; returns the lowest step energy (the reference reports
        the minimum, src/optimize_network.c:650-654) and the step records."""
        recs = [self.step(k) for k in range(len(self.steps))]
        e = min(r["energy"] for r in recs)
        self.energies.append(e)
        return e, recs

    def run_scheme(self, regimes):
        """execute_optScheme (src/optimize_network.c:640-660): every regime
        (a Regime, or a tuple in its field order, which is that of struct
        regime, include/optScheme.h:28-45) runs sweeps at its settings until
        two consecutive sweep energies differ by less than its e_conv or
        max_sweeps sweeps have run (execute_regime,
        src/optimize_network.c:395-420); a regime's energy is its last
        sweep's, and the result is the lowest regime energy."""
        self.regime_energies = []
        for reg in regimes:
            reg = Regime(*reg)
            if reg.max_sweeps < 1:
                raise ValueError("a regime with no sweeps")
            self.min_dim, self.max_dim = int(reg.min_dim), int(reg.max_dim)
            self.truncerr, self.tol = reg.truncerr, reg.davidson_tol
            if hasattr(self.solver, "max_its"):
                self.solver.max_its = int(reg.davidson_max_its)
            for i in range(int(reg.max_sweeps)):
                self.sweep()
                if i > 0 and abs(self.energies[-1] - self.energies[-2]) < reg.e_conv:
                    break
            self.regime_energies.append(self.energies[-1])
        return min(self.regime_energies)

    def run(self, sweeps, e_conv=0.0):
        """Up to `sweeps` sweeps, stopping when two consecutive sweep energies
        differ by less than e_conv (src/optimize_network.c:410-415)."""
        for i in range(sweeps):
            self.sweep()
            if i > 0 and abs(self.energies[-1] - self.energies[-2]) < e_conv:
                break
        return min(self.energies)
