"""ORACLE -- test infrastructure, not product code.

ctypes binding of oracle/liboracle.so (the plain-C restatement of the
reference's Heff path, see oracle.h) and of oracle/_ref/libref_sym.so (the
reference's own symmetry sources compiled unchanged).  Only tests/, the smoke
check and bench.py's CPU-baseline leg import this package, as the checker.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "liboracle.so")
REFPATH = os.path.join(_HERE, "_ref", "libref_sym.so")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError("oracle not built: run oracle/build.py")
        from t3ns_amd._lib import (AppendDesc, BUpdateDesc, HeffDesc, MakesiteDesc, PUpdateDesc,
                                   QRDesc, QRView, RopsDesc, SiteDesc, SplitDesc, SplitView)
        L = C.CDLL(LIBPATH)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        i64 = C.POINTER(C.c_int64)
        sig = {
            "oracle_wigner6j": (C.c_double, [C.c_int] * 6),
            "oracle_wigner9j": (C.c_double, [C.c_int] * 9),
            "oracle_prefactor_add_P_operator": (C.c_double, [ip, C.c_int, ip, C.c_int]),
            "oracle_prefactor_combine_MPOs": (C.c_double, [ip, ip, ip, C.c_int, C.c_int, C.c_int]),
            "oracle_prefactor_pAppend": (C.c_double, [ip, C.c_int, ip, C.c_int]),
            "oracle_prefactor_adjoint": (C.c_double, [ip, C.c_char, ip, C.c_int]),
            "oracle_prefactor_pUpdate": (C.c_double, [ip, C.c_int, ip, C.c_int]),
            "oracle_prefactor_mirror_coupling": (C.c_double, [ip, ip, C.c_int]),
            "oracle_prefactor_bUpdate": (C.c_double, [ip, C.c_int, ip, C.c_int]),
            "oracle_dgemm_batched": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                            C.c_double, dp, C.c_int, C.c_int64, dp, C.c_int,
                                            C.c_int64, C.c_double, dp, C.c_int, C.c_int64]),
            "oracle_heff_init": (C.c_int, [C.POINTER(P), C.POINTER(HeffDesc), C.POINTER(RopsDesc),
                                           C.POINTER(SiteDesc)]),
            "oracle_heff_size": (C.c_int64, [P]),
            "oracle_heff_matvec": (None, [P, dp, dp]),
            "oracle_heff_matvec_blocks": (None, [P, dp, dp, C.c_int, C.c_int, C.c_int]),
            "oracle_heff_block_stats": (None, [P, i64, i64, i64, dp, dp]),
            "oracle_heff_diagonal": (None, [P, dp]),
            "oracle_heff_plan_stats": (None, [P, i64, i64, i64, dp]),
            "oracle_heff_plan_table": (C.c_int64, [P, ip, dp]),
            "oracle_heff_destroy": (None, [P]),
            "oracle_append_phys": (C.c_int64, [C.POINTER(AppendDesc), C.POINTER(RopsDesc), i64, dp]),
            "oracle_update_physical": (C.c_int64, [C.POINTER(PUpdateDesc), C.POINTER(RopsDesc),
                                                   i64, dp]),
            "oracle_make_site_tensor": (C.c_int, [C.POINTER(MakesiteDesc), dp]),
            "oracle_update_branching": (C.c_int64, [C.POINTER(BUpdateDesc), C.POINTER(RopsDesc),
                                                    i64, dp]),
            "oracle_set_bupdate_order": (None, [C.c_int]),
            "oracle_split_site": (C.c_int, [C.POINTER(SplitDesc), C.POINTER(SplitView)]),
            "oracle_split_free": (None, [C.POINTER(SplitView)]),
            "oracle_qr_step": (C.c_int, [C.POINTER(QRDesc), C.POINTER(QRView)]),
            "oracle_qr_free": (None, [C.POINTER(QRView)]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        L.oracle_use_blas.restype = C.c_int
        L.oracle_use_blas.argtypes = [C.c_char_p]
        L.oracle_blas_active.restype = C.c_int
        _lib = L
    return _lib


def _openblas_path():
    """scipy's bundled OpenBLAS (CBLAS symbols with the scipy_ prefix), the
    GEMM the oracle uses in place of the reference's cblas_dgemm; None if
    scipy is not installed."""
    import glob
    try:
        import scipy
    except ImportError:
        return None
    libs = glob.glob(os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs",
                                  "libscipy_openblas*.so"))
    return libs[0] if libs else None


def use_blas():
    """Runs the oracle's GEMMs on OpenBLAS from here on (the CPU baseline:
    the reference calls cblas_dgemm).  The checker itself keeps the plain
    loop nest unless this is called.  Returns whether BLAS is in use."""
    path = _openblas_path()
    if path and not lib().oracle_blas_active():
        lib().oracle_use_blas(path.encode())
    return blas_active()


def blas_active():
    """True when the oracle's GEMMs run on OpenBLAS."""
    return bool(lib().oracle_blas_active())


def ref():
    """The reference's compiled symmetry library, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REFPATH):
        _ref = C.CDLL(REFPATH)
    return _ref


class OracleHeff:
    def __init__(self, desc, ops, site):
        h = C.c_void_p()
        rc = lib().oracle_heff_init(C.byref(h), C.byref(desc), ops, C.byref(site))
        if rc != 0:
            raise RuntimeError("oracle_heff_init failed")
        self._h = h
        self.size = int(lib().oracle_heff_size(h))
        self.nrblocks = int(site.nrblocks)

    def matvec(self, vec):
        vec = np.ascontiguousarray(vec, dtype=np.float64)
        out = np.empty(self.size)
        lib().oracle_heff_matvec(self._h, vec.ctypes.data_as(C.POINTER(C.c_double)),
                                 out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def matvec_blocks(self, vec, out, first, last, stride=1):
        """Fills only output blocks first, first+stride, ... < last of out."""
        lib().oracle_heff_matvec_blocks(self._h, vec.ctypes.data_as(C.POINTER(C.c_double)),
                                        out.ctypes.data_as(C.POINTER(C.c_double)),
                                        int(first), int(last), int(stride))

    def block_flops(self):
        n = self.nrblocks
        out = np.zeros(n)
        lib().oracle_heff_block_stats(self._h, None, None, None, None,
                                      out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def diagonal(self):
        out = np.empty(self.size)
        lib().oracle_heff_diagonal(self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def plan_stats(self):
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        lib().oracle_heff_plan_stats(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return {"npairs": a.value, "nentries": b.value, "nterms": c.value, "flops": d.value}

    def plan_table(self):
        n = lib().oracle_heff_plan_table(self._h, None, None)
        ints = np.zeros((max(n, 1), 7), dtype=np.int32)
        prefs = np.zeros(max(n, 1))
        lib().oracle_heff_plan_table(self._h, ints.ctypes.data_as(C.POINTER(C.c_int)),
                                     prefs.ctypes.data_as(C.POINTER(C.c_double)))
        return ints[:n], prefs[:n]

    def close(self):
        if self._h:
            lib().oracle_heff_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def append_phys(desc, old):
    """Oracle rOperators_append_phys: [(beginblock, elements)] per new operator."""
    import ctypes as _C
    nops = 0
    for i in range(desc.nr_instr):
        nops = max(nops, desc.instr[3 * i + 2] + 1)
    nbb = 0
    for k in range(nops):
        h = desc.hss_of_new[k]
        nbb += desc.begin_blocks_of_hss[h + 1] - desc.begin_blocks_of_hss[h] + 1
    total = lib().oracle_append_phys(_C.byref(desc), _C.byref(old), None, None)
    bb = np.zeros(max(nbb, 1), dtype=np.int64)
    tel = np.zeros(max(total, 1))
    lib().oracle_append_phys(_C.byref(desc), _C.byref(old),
                             bb.ctypes.data_as(_C.POINTER(_C.c_int64)),
                             tel.ctypes.data_as(_C.POINTER(_C.c_double)))
    out, pb, pt = [], 0, 0
    for k in range(nops):
        h = desc.hss_of_new[k]
        nb = desc.begin_blocks_of_hss[h + 1] - desc.begin_blocks_of_hss[h]
        b = bb[pb:pb + nb + 1].copy()
        out.append((b, tel[pt:pt + int(b[-1])].copy()))
        pb += nb + 1
        pt += int(b[-1])
    return out


def update_physical(desc, pops):
    """Oracle update_rOperators_physical: [(beginblock, elements)] per operator."""
    import ctypes as _C
    nbb = 0
    for k in range(pops.nrops):
        h = pops.hss_of_ops[k]
        nbb += desc.begin_blocks_of_hss[h + 1] - desc.begin_blocks_of_hss[h] + 1
    total = lib().oracle_update_physical(_C.byref(desc), _C.byref(pops), None, None)
    bb = np.zeros(max(nbb, 1), dtype=np.int64)
    tel = np.zeros(max(total, 1))
    lib().oracle_update_physical(_C.byref(desc), _C.byref(pops),
                                 bb.ctypes.data_as(_C.POINTER(_C.c_int64)),
                                 tel.ctypes.data_as(_C.POINTER(_C.c_double)))
    out, pb, pt = [], 0, 0
    for k in range(pops.nrops):
        h = pops.hss_of_ops[k]
        nb = desc.begin_blocks_of_hss[h + 1] - desc.begin_blocks_of_hss[h]
        b = bb[pb:pb + nb + 1].copy()
        out.append((b, tel[pt:pt + int(b[-1])].copy()))
        pb += nb + 1
        pt += int(b[-1])
    return out


def make_site_tensor(desc):
    """Oracle makesiteTensor: the multisite elements."""
    n = int(desc.beginblock[desc.nrblocks]) if desc.nrblocks else 0
    out = np.zeros(max(n, 1))
    rc = lib().oracle_make_site_tensor(C.byref(desc), out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise RuntimeError("oracle_make_site_tensor: sites not bound to the inner site")
    return out[:n]


def update_branching(desc, ops, order=-1):
    """Oracle update_rOperators_branching: [(beginblock, elements)] per new
    operator.  order 0..5 forces one contraction order for every block
    (default: the cheapest, as the reference)."""
    lib().oracle_set_bupdate_order(int(order))
    nops = 0
    for i in range(desc.nr_instr):
        nops = max(nops, desc.instr[3 * i + 2] + 1)
    nbb = 0
    for k in range(nops):
        h = desc.hss_of_new[k]
        nbb += desc.begin_blocks_of_hss[h + 1] - desc.begin_blocks_of_hss[h] + 1
    total = lib().oracle_update_branching(C.byref(desc), ops, None, None)
    bb = np.zeros(max(nbb, 1), dtype=np.int64)
    tel = np.zeros(max(total, 1))
    lib().oracle_update_branching(C.byref(desc), ops, bb.ctypes.data_as(C.POINTER(C.c_int64)),
                                  tel.ctypes.data_as(C.POINTER(C.c_double)))
    lib().oracle_set_bupdate_order(-1)
    out, pb, pt = [], 0, 0
    for k in range(nops):
        h = desc.hss_of_new[k]
        nb = desc.begin_blocks_of_hss[h + 1] - desc.begin_blocks_of_hss[h]
        b = bb[pb:pb + nb + 1].copy()
        out.append((b, tel[pt:pt + int(b[-1])].copy()))
        pb += nb + 1
        pt += int(b[-1])
    return out


def split_site(desc):
    """Oracle split_of_site: the result dict of t3ns_amd.split_site."""
    from t3ns_amd._lib import SplitView, split_view_dict
    v = SplitView()
    if lib().oracle_split_site(C.byref(desc), C.byref(v)) != 0:
        raise RuntimeError("oracle_split_site: site not cut by one bond")
    try:
        return split_view_dict(v)
    finally:
        lib().oracle_split_free(C.byref(v))


def qr_step(desc):
    """Oracle qr_step: the result dict of t3ns_amd.qr_step."""
    from t3ns_amd._lib import QRView, qr_view_dict
    v = QRView()
    if lib().oracle_qr_step(C.byref(desc), C.byref(v)) != 0:
        raise RuntimeError("oracle_qr_step: the tensors share no bond")
    try:
        return qr_view_dict(v)
    finally:
        lib().oracle_qr_free(C.byref(v))
