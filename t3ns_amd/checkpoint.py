"""Checkpoints of a sweep in the reference's HDF5 layout (T3NScalc.h5,
src/io_to_disk.c:446-487) through t3ns_ckpt_* (include/t3ns_host.h):
the network, the bookkeeper's current bond sectors, the Hamiltonian and one
site tensor per network site.  Operators are not stored; a restored sweep
rebuilds them from the tensors, as its environment does at start-up."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .heff import Context

# enum hamtypes (include/hamiltonian.h:30)
HAM_TYPES = {1: "QC", 2: "NN_HUBBARD"}


def write_checkpoint(path, ctx, tensors):
    """Writes `ctx` and `tensors` ({site: SiteData}, one per network site)
    to `path` (write_to_disk, src/io_to_disk.c:446-466, with the file path
    given in full)."""
    n = ctx.info()["sites"]
    if sorted(tensors) != list(range(n)):
        raise ValueError("one site tensor per network site is needed")
    keep = [tensors[s] for s in range(n)]
    nrb = (C.c_int * n)(*[t.nrblocks for t in keep])
    qn = (_lib.c_i64_p * n)(*[t.qn.ctypes.data_as(_lib.c_i64_p) for t in keep])
    bb = (_lib.c_i64_p * n)(*[t.bb.ctypes.data_as(_lib.c_i64_p) for t in keep])
    tel = (_lib.c_dbl_p * n)(*[t.tel.ctypes.data_as(_lib.c_dbl_p) for t in keep])
    check(lib().t3ns_ckpt_write(ctx.handle, str(path).encode(), n, nrb, qn, bb, tel),
          "t3ns_ckpt_write")


def read_checkpoint(path):
    """Reads `path` (read_from_disk, src/io_to_disk.c:468-487): returns a
    new Context (network, symmetries, Hamiltonian, bond sectors as stored),
    the site tensors {site: SiteData} and a dict of what else the file
    holds (model name, operator groups present)."""
    from .sweep import SiteData
    k = C.c_void_p()
    check(lib().t3ns_ckpt_open(C.byref(k), str(path).encode()), "t3ns_ckpt_open")
    try:
        ns, nb, ht, nro = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().t3ns_ckpt_info(k, C.byref(ns), C.byref(nb), C.byref(ht), C.byref(nro)),
              "t3ns_ckpt_info")
        tensors = {}
        for s in range(ns.value):
            nrs, nrb = C.c_int(), C.c_int()
            qn, bb, tel = _lib.c_i64_p(), _lib.c_i64_p(), _lib.c_dbl_p()
            check(lib().t3ns_ckpt_site(k, s, C.byref(nrs), C.byref(nrb), C.byref(qn),
                                       C.byref(bb), C.byref(tel)), "t3ns_ckpt_site")
            if nrs.value != 1:
                raise ValueError("site %d holds a %d-site tensor; single-site tensors are "
                                 "expected" % (s, nrs.value))
            n = nrb.value
            q = np.ctypeslib.as_array(qn, (n,)).copy() if n else np.zeros(0, np.int64)
            b = np.ctypeslib.as_array(bb, (n + 1,)).copy()
            m = int(b[-1])
            t = np.ctypeslib.as_array(tel, (m,)).copy() if m else np.zeros(0)
            tensors[s] = SiteData(q, b, t)
        h = C.c_void_p()
        check(lib().t3ns_ckpt_context(k, C.byref(h)), "t3ns_ckpt_context")
        ctx = Context._from_handle(h)
    finally:
        lib().t3ns_ckpt_close(k)
    return ctx, tensors, {"model": HAM_TYPES.get(ht.value, ht.value), "nr_bonds": nb.value,
                          "operator_groups": nro.value}
