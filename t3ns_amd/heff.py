"""Host-side mirror of the reference's Heff interface on top of the C ABI.

Reference API (include/Heff.h:166-192)   ->  here
    init_Heffdata(data, Operators, site)  ->  Heff(step_or_descs, precision)
    matvecT3NS(vec, result, vdata)        ->  matvecT3NS(vec, result, heff)
    make_diagonal(data)                   ->  make_diagonal(heff)
    destroy_Heffdata(data)                ->  destroy_Heffdata(heff) / Heff.close()
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (FP32, FP64, AppendDesc, BUpdateDesc, HeffDesc, MakesiteDesc, PUpdateDesc, QRDesc,
                   QRView, RopsDesc, SiteDesc, SplitDesc, SplitView, Work, check, lib, qr_view_dict,
                   split_view_dict)


class Context:
    """Solve context: network, model, symmetry sectors of every bond.

    Replaces make_network + readinteraction + preparebookkeeper of the
    reference's tests (tests/test1.c.in:31-54)."""

    def __init__(self, netw_text, model, params, symmetries, target, max_dim, minocc=2,
                 fcidump=None):
        """model "NN_HUBBARD" takes params (t, U); model "QC" takes the FCIDUMP
        text in `fcidump` (params unused)."""
        sgs = [_lib.SYMGROUPS[s] if isinstance(s, str) else int(s) for s in symmetries]
        ts = [int(t) for t in target]
        csgs, cts = (C.c_int * len(sgs))(*sgs), (C.c_int * len(ts))(*ts)
        h = C.c_void_p()
        if model == "QC":
            if fcidump is None:
                raise ValueError("model QC needs the FCIDUMP text")
            check(lib().t3ns_ctx_create_qc(C.byref(h), netw_text.encode(), fcidump.encode(),
                                           csgs, cts, len(sgs), int(max_dim), int(minocc)),
                  "t3ns_ctx_create_qc")
        else:
            par = (C.c_double * max(len(params), 1))(*params)
            check(lib().t3ns_ctx_create(C.byref(h), netw_text.encode(), model.encode(), par,
                                        len(params), csgs, cts, len(sgs), int(max_dim),
                                        int(minocc)),
                  "t3ns_ctx_create")
        self._h = h

    @classmethod
    def _from_handle(cls, h):
        obj = cls.__new__(cls)
        obj._h = h
        return obj

    @property
    def handle(self):
        return self._h

    def info(self):
        nb, ns, nps, nsym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        sgs = (C.c_int * 5)()
        check(lib().t3ns_ctx_info(self._h, C.byref(nb), C.byref(ns), C.byref(nps), C.byref(nsym), sgs))
        return {"nr_bonds": nb.value, "sites": ns.value, "psites": nps.value,
                "sgs": list(sgs)[:nsym.value]}

    def bond_dims(self, bond):
        n, tot = C.c_int(), C.c_int()
        check(lib().t3ns_ctx_bond(self._h, bond, C.byref(n), C.byref(tot), None, 0))
        buf = (C.c_int * max(n.value, 1))()
        check(lib().t3ns_ctx_bond(self._h, bond, C.byref(n), C.byref(tot), buf, n.value))
        return list(buf)[:n.value]

    def sweep_steps(self, maxsites):
        """Site lists of the optimisation steps of one sweep."""
        n = lib().t3ns_ctx_sweep_steps(self._h, int(maxsites), None, None, 0)
        if n < 0:
            check(n, "t3ns_ctx_sweep_steps")
        sites = (C.c_int * (4 * max(n, 1)))()
        ns = (C.c_int * max(n, 1))()
        lib().t3ns_ctx_sweep_steps(self._h, int(maxsites), sites, ns, n)
        return [list(sites[4 * i:4 * i + ns[i]]) for i in range(n)]

    def qr_problem(self, site_a, site_c, seed=0):
        """qr_step descriptor: seeded one-site tensors of neighbouring sites
        site_a (decomposed) and site_c (the next orthogonality centre)."""
        d = QRDesc()
        check(lib().t3ns_ctx_qr_desc(self._h, int(site_a), int(site_c), int(seed), C.byref(d)),
              "t3ns_ctx_qr_desc")
        return d

    def bupdate_problem(self, site, update_case, seed=0):
        """(update descriptor, old operators [2]) of update_rOperators_branching
        at branching site `site` for new operators on its bond `update_case`."""
        d, o = BUpdateDesc(), (RopsDesc * 2)()
        check(lib().t3ns_ctx_bupdate_desc(self._h, int(site), int(update_case), int(seed),
                                          C.byref(d), o), "t3ns_ctx_bupdate_desc")
        return d, o

    def bond_sites(self, bond):
        """(left site, right site) of bond `bond`, -1 at a border."""
        out = (C.c_int * 2)()
        check(lib().t3ns_ctx_bond_sites(self._h, int(bond), out), "t3ns_ctx_bond_sites")
        return out[0], out[1]

    def site_bonds(self, site):
        """([three bonds], is physical) of site `site`."""
        out, ph = (C.c_int * 3)(), C.c_int()
        check(lib().t3ns_ctx_site_bonds(self._h, int(site), out, C.byref(ph)), "t3ns_ctx_site_bonds")
        return list(out), bool(ph.value)

    def fill_seeded(self, on):
        """Whether problem builders fill seeded elements (default) or give
        structure only with zero elements."""
        check(lib().t3ns_ctx_fill_seeded(self._h, int(bool(on))), "t3ns_ctx_fill_seeded")

    def set_bond(self, bond, irreps, dims):
        """Replaces the sectors of bond `bond`: irreps [n][MAX_SYMMETRIES],
        dims [n] (as a decomposition leaves them)."""
        ir = np.ascontiguousarray(irreps, dtype=np.int32).reshape(-1)
        dm = np.ascontiguousarray(dims, dtype=np.int32)
        if ir.size != dm.size * _lib.MAX_SYMMETRIES:
            raise ValueError("set_bond: irreps and dims disagree")
        check(lib().t3ns_ctx_set_bond(self._h, int(bond), int(dm.size),
                                      ir.ctypes.data_as(_lib.c_int_p), dm.ctypes.data_as(_lib.c_int_p)),
              "t3ns_ctx_set_bond")

    def site_tensor(self, site, seed=0):
        """One-site tensor of `site` over the current sectors with seeded
        N(0,1) elements: (qnumbers, beginblock, elements) as numpy copies."""
        n = C.c_int()
        qn, bb, tel = _lib.c_i64_p(), _lib.c_i64_p(), _lib.c_dbl_p()
        check(lib().t3ns_ctx_site_tensor(self._h, int(site), int(seed), C.byref(n), C.byref(qn),
                                         C.byref(bb), C.byref(tel)), "t3ns_ctx_site_tensor")
        nb = n.value
        q = np.ctypeslib.as_array(qn, shape=(max(nb, 1),))[:nb].copy()
        b = np.ctypeslib.as_array(bb, shape=(nb + 1,)).copy()
        t = np.ctypeslib.as_array(tel, shape=(max(int(b[-1]), 1),))[:int(b[-1])].copy()
        return q, b, t

    def append_problem_at(self, bond, is_left, seed=0):
        """(append descriptor, seeded old operators) of appending the
        physical site on side `is_left` of bond `bond` to the operators of
        that site's far bond."""
        d, o = AppendDesc(), RopsDesc()
        check(lib().t3ns_ctx_append_desc(self._h, int(bond), int(is_left), int(seed), C.byref(d),
                                         C.byref(o)), "t3ns_ctx_append_desc")
        return d, o

    def pupdate_problem_at(self, bond, is_left, seed=0):
        """update_rOperators_physical descriptor on bond `bond` (side
        `is_left`) with a seeded one-site tensor over the context's sectors."""
        d = PUpdateDesc()
        check(lib().t3ns_ctx_pupdate_desc(self._h, int(bond), int(is_left), int(seed), C.byref(d)),
              "t3ns_ctx_pupdate_desc")
        return d

    def close(self):
        if self._h:
            lib().t3ns_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Step:
    """Inputs of one Heff (multisite tensor + renormalised operators)."""

    def __init__(self, ctx, sites, seed=0):
        h = C.c_void_p()
        arr = (C.c_int * len(sites))(*sites)
        check(lib().t3ns_step_random(C.byref(h), ctx.handle, arr, len(sites), int(seed)),
              "t3ns_step_random")
        self._h = h
        self.ctx = ctx  # keeps the context alive

    @property
    def size(self):
        return int(lib().t3ns_step_size(self._h))

    def descriptors(self):
        d, o, s = HeffDesc(), (RopsDesc * 3)(), SiteDesc()
        check(lib().t3ns_step_heff_desc(self._h, C.byref(d), o, C.byref(s)), "t3ns_step_heff_desc")
        return d, o, s

    def append_problem(self, leg, seed=0):
        """(append descriptor, old operators) behind physical leg `leg`."""
        d, o = AppendDesc(), RopsDesc()
        check(lib().t3ns_step_append_desc(self._h, int(leg), int(seed), C.byref(d), C.byref(o)),
              "t3ns_step_append_desc")
        return d, o

    def pupdate_problem(self, leg, seed=0, maxdim=8):
        """(update descriptor, physical operators) behind physical leg `leg`:
        the leg's bond truncated to a seeded subset of its internal sectors
        (dims up to `maxdim`) and a seeded one-site tensor on its site."""
        d, o = PUpdateDesc(), RopsDesc()
        check(lib().t3ns_step_pupdate_desc(self._h, int(leg), int(seed), int(maxdim), C.byref(d),
                                           C.byref(o)), "t3ns_step_pupdate_desc")
        return d, o

    def split_problem(self, siteV, maxD, minD=0, truncerr=0.0, trunc_type="E"):
        """decompose_siteTensor split of site index `siteV` off the step's
        (normalised) multisite tensor."""
        d = SplitDesc()
        check(lib().t3ns_step_split_desc(self._h, int(siteV), int(minD), int(maxD), float(truncerr),
                                         trunc_type.encode(), C.byref(d)), "t3ns_step_split_desc")
        return d

    def makesite_problem(self, seed=0):
        """makesiteTensor descriptor: seeded one-site tensors of the step's
        sites and the step's multisite structure."""
        d = MakesiteDesc()
        check(lib().t3ns_step_makesite_desc(self._h, int(seed), C.byref(d)), "t3ns_step_makesite_desc")
        return d

    def close(self):
        if self._h:
            lib().t3ns_step_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_table(desc, ops, site):
    """Host-built dispatch table: (ints[n,7], prefactors[n], flops)."""
    fl = C.c_double()
    n = lib().t3ns_plan_table(C.byref(desc), ops, C.byref(site), None, None, C.byref(fl))
    if n < 0:
        check(int(n), "t3ns_plan_table")
    ints = np.zeros((max(n, 1), 7), dtype=np.int32)
    prefs = np.zeros(max(n, 1), dtype=np.float64)
    lib().t3ns_plan_table(C.byref(desc), ops, C.byref(site),
                          ints.ctypes.data_as(_lib.c_int_p), prefs.ctypes.data_as(_lib.c_dbl_p), None)
    return ints[:n], prefs[:n], fl.value


class ROps:
    """Renormalised operators owned by the library (host copy)."""

    def __init__(self, handle):
        self._h = handle
        self.desc = RopsDesc()
        check(lib().t3ns_rops_view(self._h, C.byref(self.desc)), "t3ns_rops_view")

    def blocks(self):
        """[(beginblock array, elements array)] per operator."""
        d = self.desc
        out = []
        for k in range(d.nrops):
            h = d.hss_of_ops[k]
            nb = d.begin_blocks_of_hss[h + 1] - d.begin_blocks_of_hss[h]
            bb = np.ctypeslib.as_array(d.beginblock[k], shape=(nb + 1,)).copy()
            n = int(bb[-1])
            tel = np.ctypeslib.as_array(d.tel[k], shape=(max(n, 1),))[:n].copy() if n else np.zeros(0)
            out.append((bb, tel))
        return out

    def work(self):
        """(flops, bytes, kernel ms) of the device execution that made these."""
        f, b, ms = C.c_double(), C.c_double(), C.c_double()
        check(lib().t3ns_rops_work(self._h, C.byref(f), C.byref(b), C.byref(ms)), "t3ns_rops_work")
        return f.value, b.value, ms.value

    def close(self):
        if self._h:
            lib().t3ns_rops_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def append_phys(desc, old):
    """rOperators_append_phys (src/rOperators_pUpdate.c:665-673) on the device."""
    h = C.c_void_p()
    check(lib().t3ns_append_phys(C.byref(desc), C.byref(old), C.byref(h)), "t3ns_append_phys")
    return ROps(h)


def update_physical(desc, pops):
    """update_rOperators_physical (src/rOperators_pUpdate.c:364-387) on the
    device; `pops` is a RopsDesc or an ROps."""
    if isinstance(pops, ROps):
        pops = pops.desc
    h = C.c_void_p()
    check(lib().t3ns_update_physical(C.byref(desc), C.byref(pops), C.byref(h)),
          "t3ns_update_physical")
    return ROps(h)


def update_branching(desc, ops):
    """update_rOperators_branching (src/rOperators_bUpdate.c:229-248) on the
    device; `ops` is the array of the two old operator descriptors."""
    h = C.c_void_p()
    check(lib().t3ns_update_branching(C.byref(desc), ops, C.byref(h)), "t3ns_update_branching")
    return ROps(h)


def split_site(desc):
    """One split of decompose_siteTensor (src/siteTensor_decompose.c:
    1403-1444) on the device: (result dict, (flops, bytes, kernel ms))."""
    h = C.c_void_p()
    w = Work()
    check(lib().t3ns_split_site(C.byref(desc), C.byref(h), C.byref(w)), "t3ns_split_site")
    try:
        v = SplitView()
        check(lib().t3ns_split_get(h, C.byref(v)), "t3ns_split_get")
        out = split_view_dict(v)
    finally:
        lib().t3ns_split_destroy(h)
    return out, (w.flops, w.bytes, w.kernel_ms)


def qr_step(desc):
    """qr_step (src/siteTensor_decompose.c:1587-1630) on the device:
    (result dict, (flops, bytes, kernel ms))."""
    h = C.c_void_p()
    w = Work()
    check(lib().t3ns_qr_step(C.byref(desc), C.byref(h), C.byref(w)), "t3ns_qr_step")
    try:
        v = QRView()
        check(lib().t3ns_qr_get(h, C.byref(v)), "t3ns_qr_get")
        out = qr_view_dict(v)
    finally:
        lib().t3ns_qr_destroy(h)
    return out, (w.flops, w.bytes, w.kernel_ms)


def make_site_tensor(desc):
    """makesiteTensor (src/siteTensor_init.c:785-800) on the device:
    (multisite elements, (flops, bytes, kernel ms))."""
    n = int(desc.beginblock[desc.nrblocks]) if desc.nrblocks else 0
    out = np.zeros(max(n, 1))
    w = Work()
    check(lib().t3ns_make_site_tensor(C.byref(desc), out.ctypes.data_as(_lib.c_dbl_p), C.byref(w)),
          "t3ns_make_site_tensor")
    return out[:n], (w.flops, w.bytes, w.kernel_ms)


def plan_block_dims(desc, ops, site):
    """Leg dimensions [nrblocks, 3] of the multisite tensor's blocks."""
    dims = np.zeros((max(site.nrblocks, 1), 3), dtype=np.int32)
    n = lib().t3ns_plan_block_dims(C.byref(desc), ops, C.byref(site),
                                   dims.ctypes.data_as(_lib.c_int_p))
    if n < 0:
        check(n, "t3ns_plan_block_dims")
    return dims[:n]


def _vec(a, n, what, writable=False):
    """`a` as a C-contiguous float64 array of shape (n,); raises ValueError
    otherwise (the C side reads or writes exactly n doubles).  A writable
    output must already have that form (no silent copy)."""
    if writable:
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.shape == (n,)
                and a.flags.c_contiguous and a.flags.writeable):
            raise ValueError("%s must be a writable C-contiguous float64 array of %d elements"
                             % (what, n))
        return a
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != (n,):
        raise ValueError("%s has shape %s, the operator needs (%d,)" % (what, a.shape, n))
    return a


class Heff:
    """init_Heffdata (include/Heff.h:166) on the device."""

    def __init__(self, desc, ops, site, precision=FP64, keepalive=None):
        h = C.c_void_p()
        check(lib().t3ns_heff_init(C.byref(h), C.byref(desc), ops, C.byref(site), int(precision)),
              "t3ns_heff_init")
        self._h = h
        self._keep = keepalive
        self.size = int(lib().t3ns_heff_size(h))

    @classmethod
    def from_step(cls, step, precision=FP64):
        d, o, s = step.descriptors()
        return cls(d, o, s, precision, keepalive=step)

    def derive_fp32(self):
        """An FP32 Heff on this FP64 object's plan and operator arenas,
        copied and converted on the device (t3ns_heff_derive): only its own
        work lists are built on the host.  Independent of this object
        afterwards."""
        h = C.c_void_p()
        check(lib().t3ns_heff_derive(C.byref(h), self._h, int(FP32)), "t3ns_heff_derive")
        out = Heff.__new__(Heff)
        out._h = h
        out._keep = self._keep
        out.size = int(lib().t3ns_heff_size(h))
        return out

    @property
    def handle(self):
        return self._h

    def matvec(self, vec, out=None):
        vec = _vec(vec, self.size, "vector")
        if out is None:
            out = np.empty(self.size, dtype=np.float64)
        else:
            out = _vec(out, self.size, "out", writable=True)
        check(lib().t3ns_heff_matvec(self._h, vec.ctypes.data_as(_lib.c_dbl_p),
                                     out.ctypes.data_as(_lib.c_dbl_p)), "t3ns_heff_matvec")
        return out

    def matvec_device(self, dvec_ptr, dres_ptr, stream_ptr=None):
        check(lib().t3ns_heff_matvec_device(self._h, C.c_void_p(dvec_ptr), C.c_void_p(dres_ptr),
                                            C.c_void_p(stream_ptr or 0)), "t3ns_heff_matvec_device")

    def use_graphs(self, on=True):
        """Replay matvecs on a non-default stream from captured HIP graphs
        (t3ns_heff_use_graphs)."""
        check(lib().t3ns_heff_use_graphs(self._h, int(bool(on))), "t3ns_heff_use_graphs")

    def diagonal(self):
        out = np.empty(self.size, dtype=np.float64)
        check(lib().t3ns_heff_diagonal(self._h, out.ctypes.data_as(_lib.c_dbl_p)), "t3ns_heff_diagonal")
        return out

    def stats(self):
        nb, npair, nt, fl = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        check(lib().t3ns_heff_stats(self._h, C.byref(nb), C.byref(npair), C.byref(nt), C.byref(fl)))
        return {"nblocks": nb.value, "npairs": npair.value, "nterms": nt.value, "flops": fl.value,
                "size": self.size}

    def davidson(self, guess, max_vecs=30, keep_deflate=2, tol=1e-8, max_its=100,
                 diagonal=None):
        """Lowest eigenpair (davidson, include/davidson.h:56-60): returns
        (energy, eigenvector, iterations, status) with status 0 converged,
        1 stopped at max_its."""
        x = np.array(_vec(guess, self.size, "guess"), copy=True)
        e, its = C.c_double(), C.c_int()
        dg = None if diagonal is None else _vec(diagonal, self.size, "diagonal")
        rc = lib().t3ns_davidson(self._h, x.ctypes.data_as(_lib.c_dbl_p), C.byref(e),
                                 int(max_vecs), int(keep_deflate), float(tol), int(max_its),
                                 None if dg is None else dg.ctypes.data_as(_lib.c_dbl_p),
                                 C.byref(its))
        if rc < 0:
            check(rc, "t3ns_davidson")
        return e.value, x, its.value, rc

    TUNING = ("chunk_steps", "fuse_over", "fuse", "sched", "wpc")

    def autotune(self, max_trials=6):
        """Times matvecs under up to max_trials work-list settings and keeps
        the fastest (t3ns_heff_autotune); a block-size histogram tuned
        before in this process reuses its setting untimed.  Returns the
        setting with the timings."""
        buf = (C.c_double * 9)()
        check(lib().t3ns_heff_autotune(self._h, int(max_trials), buf, 9), "t3ns_heff_autotune")
        v = list(buf)
        out = dict(zip(self.TUNING[:4], v[:4]))
        out["wpc"] = v[8]
        out.update({"ms": v[4], "baseline_ms": v[5], "trials": int(v[6]), "cached": bool(v[7])})
        return out

    def tuning(self):
        buf = (C.c_double * 5)()
        check(lib().t3ns_heff_get_tuning(self._h, buf), "t3ns_heff_get_tuning")
        return dict(zip(self.TUNING, list(buf)))

    def set_tuning(self, chunk_steps=1000.0, fuse_over=2.0, fuse=1, sched=0, wpc=0):
        """wpc: workgroups per CU (1, 2 or 4; 0 picks by the object's
        structure)."""
        buf = (C.c_double * 5)(float(chunk_steps), float(fuse_over), float(fuse), float(sched),
                               float(wpc))
        check(lib().t3ns_heff_set_tuning(self._h, buf), "t3ns_heff_set_tuning")

    def kernel_stats(self):
        names = ["chunks", "runs", "pairs_global_x", "pairs_global_w", "mfma_steps",
                 "steps_A", "steps_B", "steps_C", "scratch_per_wg", "units", "split_blocks",
                 "pairs_fused", "wpc"]
        buf = (C.c_double * len(names))()
        check(lib().t3ns_heff_kernel_stats(self._h, buf, len(names)) < 0 and -1 or 0)
        return dict(zip(names, list(buf)))

    def plan_table(self):
        n = lib().t3ns_heff_plan_table(self._h, None, None)
        ints = np.zeros((max(n, 1), 7), dtype=np.int32)
        prefs = np.zeros(max(n, 1), dtype=np.float64)
        lib().t3ns_heff_plan_table(self._h, ints.ctypes.data_as(_lib.c_int_p),
                                   prefs.ctypes.data_as(_lib.c_dbl_p))
        return ints[:n], prefs[:n]

    def close(self):
        if self._h:
            lib().t3ns_heff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HeffBatch:
    """Batched matvecs of several Heff objects of one precision in one
    launch per kernel (t3ns_heff_batch_*): each result equals the object's
    own matvec bitwise."""

    def __init__(self):
        b = C.c_void_p()
        check(lib().t3ns_heff_batch_create(C.byref(b)), "t3ns_heff_batch_create")
        self._b = b

    def matvec(self, heffs, vecs):
        """Host arrays in, host arrays out (one per object)."""
        n = len(heffs)
        if len(vecs) != n:
            raise ValueError("one vector per object")
        xs = [_vec(v, h.size, "vector") for h, v in zip(heffs, vecs)]
        ys = [np.empty(h.size, dtype=np.float64) for h in heffs]
        hs = (C.c_void_p * n)(*[h.handle for h in heffs])
        xp = (_lib.c_dbl_p * n)(*[x.ctypes.data_as(_lib.c_dbl_p) for x in xs])
        yp = (_lib.c_dbl_p * n)(*[y.ctypes.data_as(_lib.c_dbl_p) for y in ys])
        check(lib().t3ns_heff_batch_matvec_host(self._b, hs, xp, yp, n),
              "t3ns_heff_batch_matvec_host")
        return ys

    def matvec_device(self, heffs, dx_ptrs, dy_ptrs, stream_ptr=0):
        """Device arrays (pointers) on `stream_ptr`, asynchronously."""
        n = len(heffs)
        hs = (C.c_void_p * n)(*[h.handle for h in heffs])
        xp = (C.c_void_p * n)(*dx_ptrs)
        yp = (C.c_void_p * n)(*dy_ptrs)
        check(lib().t3ns_heff_batch_matvec(self._b, hs, xp, yp, n, C.c_void_p(stream_ptr)),
              "t3ns_heff_batch_matvec")

    def close(self):
        if self._b:
            lib().t3ns_heff_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def clear_tune_cache():
    """Forgets every setting the autotuner recorded in this process."""
    lib().t3ns_tune_cache_clear()


def davidson_refined(h_low, h64, guess, max_vecs=30, keep_deflate=2, tol=1e-8, switch_tol=1e-4,
                     max_its=100, diagonal=None):
    """Davidson with h_low's (reduced-precision) matvec down to switch_tol,
    then FP64 matvecs down to tol.  Returns (energy, vector, its_low,
    its64, status)."""
    if h_low.size != h64.size:
        raise ValueError("the two operators differ in size")
    x = np.array(_vec(guess, h64.size, "guess"), copy=True)
    e, i1, i2 = C.c_double(), C.c_int(), C.c_int()
    dg = None if diagonal is None else _vec(diagonal, h64.size, "diagonal")
    rc = lib().t3ns_davidson_refined(h_low.handle, h64.handle, x.ctypes.data_as(_lib.c_dbl_p),
                                     C.byref(e), int(max_vecs), int(keep_deflate), float(tol),
                                     float(switch_tol), int(max_its),
                                     None if dg is None else dg.ctypes.data_as(_lib.c_dbl_p),
                                     C.byref(i1), C.byref(i2))
    if rc < 0:
        check(rc, "t3ns_davidson_refined")
    return e.value, x, i1.value, i2.value, rc


def davidson_device(size, dx_ptr, ddiag_ptr, matvec, stream_ptr=0, max_vecs=30, keep_deflate=2,
                    tol=1e-8, max_its=100):
    """Davidson on device data with a Python matvec callback
    matvec(x_ptr, y_ptr, stream_ptr) working on device pointers."""
    cb = _lib.DEV_MATVEC(lambda x, y, s, v: matvec(x, y, s))
    e, its = C.c_double(), C.c_int()
    rc = lib().t3ns_davidson_dev(int(size), C.c_void_p(dx_ptr), C.c_void_p(ddiag_ptr), cb, None,
                                 C.c_void_p(stream_ptr), C.byref(e), int(max_vecs),
                                 int(keep_deflate), float(tol), int(max_its), C.byref(its))
    if rc < 0:
        check(rc, "t3ns_davidson_dev")
    return e.value, its.value, rc


def davidson_device_noalloc(size, dx_ptr, ddiag_ptr, matvec, stream_ptr=0, max_vecs=30,
                            keep_deflate=2, tol=1e-8, max_its=100):
    """davidson_device through t3ns_davidson_dev_noalloc: no device memory
    is allocated (the solver's buffers are the library's fixed per-GPU
    array; see include/t3ns_amd.h for the size limit)."""
    cb = _lib.DEV_MATVEC(lambda x, y, s, v: matvec(x, y, s))
    e, its = C.c_double(), C.c_int()
    rc = lib().t3ns_davidson_dev_noalloc(int(size), C.c_void_p(dx_ptr), C.c_void_p(ddiag_ptr), cb,
                                         None, C.c_void_p(stream_ptr), C.byref(e), int(max_vecs),
                                         int(keep_deflate), float(tol), int(max_its), C.byref(its))
    if rc < 0:
        check(rc, "t3ns_davidson_dev_noalloc")
    return e.value, its.value, rc


# Reference-shaped entry points ------------------------------------------------

def init_Heffdata(Operators, siteObject, desc, precision=FP64):
    """init_Heffdata(data, Operators, siteObject) with the globals passed as `desc`."""
    return Heff(desc, Operators, siteObject, precision)


def matvecT3NS(vec, result, vdata):
    """matvecT3NS(vec, result, vdata): result is overwritten with Heff vec."""
    vdata.matvec(vec, out=result)


def make_diagonal(data):
    return data.diagonal()


def destroy_Heffdata(data):
    data.close()


def davidson(result, size, max_vecs, keep_deflate, davidson_tol, max_its, diagonal, vdata):
    """davidson(result, energy, size, max_vecs, keep_deflate, tol, max_its,
    diagonal, matvecT3NS, vdata): result (host array) is overwritten with the
    eigenvector; returns (status, energy)."""
    if size != vdata.size:
        raise ValueError("size does not match the operator")
    e, x, its, rc = vdata.davidson(result, max_vecs, keep_deflate, davidson_tol, max_its,
                                   diagonal)
    result[:] = x
    return rc, e
