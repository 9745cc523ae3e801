"""ctypes binding of libt3ns_amd.so (include/t3ns_amd.h, include/t3ns_host.h).

The shared library is built in-tree by t3ns_amd/build.py.  Loading fails
loudly when it is missing: there is no Python fallback for any kernel.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libt3ns_amd.so")

MAX_SYMMETRIES = 5
FP64, FP32, BF16 = 0, 1, 2
ERRORS = {0: "ok", -1: "invalid argument", -2: "device error", -3: "out of memory", -4: "solver failure"}

# symmetry group enum (include/symmetries.h:34)
Z2, U1, SU2, C1, Ci, C2, Cs, D2, C2v, C2h, D2h = range(11)
SYMGROUPS = {"Z2": Z2, "U1": U1, "SU2": SU2, "C1": C1, "Ci": Ci, "C2": C2, "Cs": Cs,
             "D2": D2, "C2v": C2v, "C2h": C2h, "D2h": D2h}

c_int_p = C.POINTER(C.c_int)
c_i64_p = C.POINTER(C.c_int64)
c_dbl_p = C.POINTER(C.c_double)


class SymsecsDesc(C.Structure):
    _fields_ = [("bond", C.c_int), ("nrSecs", C.c_int), ("irreps", c_int_p), ("dims", c_int_p)]


class RopsDesc(C.Structure):
    _fields_ = [("bond", C.c_int), ("is_left", C.c_int), ("P_operator", C.c_int),
                ("nrhss", C.c_int), ("begin_blocks_of_hss", c_int_p), ("qnumbers", c_i64_p),
                ("nrops", C.c_int), ("hss_of_ops", c_int_p),
                ("beginblock", C.POINTER(c_i64_p)), ("tel", C.POINTER(c_dbl_p))]


class SiteDesc(C.Structure):
    _fields_ = [("nrsites", C.c_int), ("sites", C.c_int * 4), ("site_is_phys", C.c_int * 4),
                ("nrblocks", C.c_int), ("qnumbers", c_i64_p), ("beginblock", c_i64_p)]


class HeffDesc(C.Structure):
    _fields_ = [("nrSyms", C.c_int), ("sgs", C.c_int * MAX_SYMMETRIES), ("mpo", SymsecsDesc),
                ("hermitian", c_int_p), ("singlet", C.POINTER(C.c_ubyte)),
                ("symarr", (SymsecsDesc * 3) * 4), ("rops_on_site", C.c_int * 3),
                ("merge_step", C.c_int), ("merge_nr", C.c_int), ("merge_instr", c_int_p),
                ("merge_pref", c_dbl_p)]


class AppendDesc(C.Structure):
    _fields_ = [("nrSyms", C.c_int), ("sgs", C.c_int * MAX_SYMMETRIES), ("is_left", C.c_int),
                ("site_ss", SymsecsDesc * 3), ("new_ss", SymsecsDesc * 3),
                ("old_ss", SymsecsDesc * 3), ("mpo", SymsecsDesc),
                ("nr_instr", C.c_int), ("instr", c_int_p), ("pref", c_dbl_p),
                ("hss_of_new", c_int_p), ("nr_siteops", C.c_int), ("siteop_hss", c_int_p),
                ("physdim", C.c_int), ("siteop_el", c_dbl_p), ("tprod_begin", c_int_p),
                ("tprods", c_int_p), ("nrhss", C.c_int), ("begin_blocks_of_hss", c_int_p),
                ("qnumbers", c_i64_p), ("hss_beginblock", C.POINTER(c_i64_p))]


class PUpdateDesc(C.Structure):
    _fields_ = [("nrSyms", C.c_int), ("sgs", C.c_int * MAX_SYMMETRIES), ("is_left", C.c_int),
                ("site_ss", SymsecsDesc * 3), ("internal_ss", SymsecsDesc), ("mpo", SymsecsDesc),
                ("T_nrblocks", C.c_int), ("T_qnumbers", c_i64_p), ("T_beginblock", c_i64_p),
                ("T_tel", c_dbl_p), ("nrhss", C.c_int), ("begin_blocks_of_hss", c_int_p),
                ("qnumbers", c_i64_p), ("hss_beginblock", C.POINTER(c_i64_p))]


class BUpdateDesc(C.Structure):
    _fields_ = [("nrSyms", C.c_int), ("sgs", C.c_int * MAX_SYMMETRIES), ("update_case", C.c_int),
                ("site_ss", SymsecsDesc * 3), ("mpo", SymsecsDesc),
                ("T_nrblocks", C.c_int), ("T_qnumbers", c_i64_p), ("T_beginblock", c_i64_p),
                ("T_tel", c_dbl_p), ("nr_instr", C.c_int), ("instr", c_int_p), ("pref", c_dbl_p),
                ("hss_of_new", c_int_p), ("tprod_begin", c_int_p), ("tprods", c_int_p),
                ("nrhss", C.c_int), ("begin_blocks_of_hss", c_int_p), ("qnumbers", c_i64_p),
                ("hss_beginblock", C.POINTER(c_i64_p))]


class SplitDesc(C.Structure):
    _fields_ = [("nrSyms", C.c_int), ("sgs", C.c_int * MAX_SYMMETRIES), ("nrsites", C.c_int),
                ("siteV", C.c_int), ("ss", (SymsecsDesc * 3) * 4), ("nrblocks", C.c_int),
                ("qnumbers", c_i64_p), ("beginblock", c_i64_p), ("tel", c_dbl_p),
                ("minD", C.c_int), ("maxD", C.c_int), ("truncerr", C.c_double),
                ("truncType", C.c_char)]


class SplitView(C.Structure):
    _fields_ = [("U_nrsites", C.c_int), ("U_nrblocks", C.c_int), ("U_qnumbers", c_i64_p),
                ("U_beginblock", c_i64_p), ("U_tel", c_dbl_p), ("V_nrblocks", C.c_int),
                ("V_qnumbers", c_i64_p), ("V_beginblock", c_i64_p), ("V_tel", c_dbl_p),
                ("bond_nrSecs", C.c_int), ("bond_irreps", c_int_p), ("bond_dims", c_int_p),
                ("nrsectors", C.c_int), ("nsing", c_int_p), ("kept", c_int_p), ("sing", c_dbl_p),
                ("norm", C.c_double * 2), ("entropy", C.c_double * 2), ("norm_U", C.c_double),
                ("sweeps", C.c_int)]


class QRDesc(C.Structure):
    _fields_ = [("nrSyms", C.c_int), ("sgs", C.c_int * MAX_SYMMETRIES), ("ssA", SymsecsDesc * 3),
                ("ssC", SymsecsDesc * 3), ("A_nrblocks", C.c_int), ("A_qnumbers", c_i64_p),
                ("A_beginblock", c_i64_p), ("A_tel", c_dbl_p), ("C_nrblocks", C.c_int),
                ("C_qnumbers", c_i64_p), ("C_beginblock", c_i64_p), ("C_tel", c_dbl_p)]


class QRView(C.Structure):
    _fields_ = [("nrblocks_Q", C.c_int), ("Q_beginblock", c_i64_p), ("Q_tel", c_dbl_p),
                ("nrblocks_B", C.c_int), ("B_beginblock", c_i64_p), ("B_tel", c_dbl_p),
                ("nrsectors", C.c_int), ("dims", c_int_p), ("nsing", c_int_p), ("sing", c_dbl_p),
                ("entropy", C.c_double)]


def qr_view_dict(v):
    """The arrays of a qr_step result view as numpy copies."""
    def arr(p, n, dt):
        return np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].astype(dt).copy() if n else np.zeros(0, dt)
    qbb = arr(v.Q_beginblock, v.nrblocks_Q + 1, np.int64)
    bbb = arr(v.B_beginblock, v.nrblocks_B + 1, np.int64)
    ns = arr(v.nsing, v.nrsectors, np.int32)
    return {"Q_beginblock": qbb, "Q_tel": arr(v.Q_tel, int(qbb[-1]), np.float64),
            "B_beginblock": bbb, "B_tel": arr(v.B_tel, int(bbb[-1]), np.float64),
            "dims": arr(v.dims, v.nrsectors, np.int32), "nsing": ns,
            "sing": arr(v.sing, int(ns.sum()), np.float64), "entropy": v.entropy}


class Work(C.Structure):
    _fields_ = [("flops", C.c_double), ("bytes", C.c_double), ("kernel_ms", C.c_double)]


class MakesiteDesc(C.Structure):
    _fields_ = [("nrsites", C.c_int), ("innersite", C.c_int),
                ("new_ss", (SymsecsDesc * 3) * 4), ("old_ss", (SymsecsDesc * 3) * 4),
                ("nrblocks", C.c_int), ("qnumbers", c_i64_p), ("beginblock", c_i64_p),
                ("site_nrblocks", C.c_int * 4), ("site_qnumbers", c_i64_p * 4),
                ("site_beginblock", c_i64_p * 4), ("site_tel", c_dbl_p * 4)]


# void (*)(const double* x, double* y, void* stream, void* vdata)
DEV_MATVEC = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p)

_lib = None


def lib():
    """Loads the in-tree library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise RuntimeError("libt3ns_amd.so is not built (run t3ns_amd/build.py): " + LIBPATH)
    L = C.CDLL(LIBPATH)
    P = C.c_void_p
    sig = {
        "t3ns_last_error": (C.c_char_p, []),
        "t3ns_device_count": (C.c_int, []),
        "t3ns_heff_init": (C.c_int, [C.POINTER(P), C.POINTER(HeffDesc), C.POINTER(RopsDesc),
                                     C.POINTER(SiteDesc), C.c_int]),
        "t3ns_heff_derive": (C.c_int, [C.POINTER(P), P, C.c_int]),
        "t3ns_heff_matvec": (C.c_int, [P, c_dbl_p, c_dbl_p]),
        "t3ns_heff_matvec_device": (C.c_int, [P, P, P, P]),
        "t3ns_heff_diagonal": (C.c_int, [P, c_dbl_p]),
        "t3ns_heff_destroy": (C.c_int, [P]),
        "t3ns_append_phys": (C.c_int, [C.POINTER(AppendDesc), C.POINTER(RopsDesc), C.POINTER(P)]),
        "t3ns_update_physical": (C.c_int, [C.POINTER(PUpdateDesc), C.POINTER(RopsDesc),
                                           C.POINTER(P)]),
        "t3ns_rops_work": (C.c_int, [P, c_dbl_p, c_dbl_p, c_dbl_p]),
        "t3ns_update_branching": (C.c_int, [C.POINTER(BUpdateDesc), C.POINTER(RopsDesc),
                                            C.POINTER(P)]),
        "t3ns_ctx_bupdate_desc": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, C.POINTER(BUpdateDesc),
                                            C.POINTER(RopsDesc)]),
        "t3ns_split_site": (C.c_int, [C.POINTER(SplitDesc), C.POINTER(P), C.POINTER(Work)]),
        "t3ns_split_get": (C.c_int, [P, C.POINTER(SplitView)]),
        "t3ns_split_destroy": (C.c_int, [P]),
        "t3ns_step_split_desc": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_char,
                                           C.POINTER(SplitDesc)]),
        "t3ns_qr_step": (C.c_int, [C.POINTER(QRDesc), C.POINTER(P), C.POINTER(Work)]),
        "t3ns_qr_get": (C.c_int, [P, C.POINTER(QRView)]),
        "t3ns_qr_destroy": (C.c_int, [P]),
        "t3ns_ctx_qr_desc": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, C.POINTER(QRDesc)]),
        "t3ns_make_site_tensor": (C.c_int, [C.POINTER(MakesiteDesc), c_dbl_p, C.POINTER(Work)]),
        "t3ns_step_makesite_desc": (C.c_int, [P, C.c_uint64, C.POINTER(MakesiteDesc)]),
        "t3ns_step_pupdate_desc": (C.c_int, [P, C.c_int, C.c_uint64, C.c_int,
                                             C.POINTER(PUpdateDesc), C.POINTER(RopsDesc)]),
        "t3ns_rops_view": (C.c_int, [P, C.POINTER(RopsDesc)]),
        "t3ns_rops_destroy": (C.c_int, [P]),
        "t3ns_step_append_desc": (C.c_int, [P, C.c_int, C.c_uint64, C.POINTER(AppendDesc),
                                            C.POINTER(RopsDesc)]),
        "t3ns_davidson": (C.c_int, [P, c_dbl_p, c_dbl_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                    c_dbl_p, c_int_p]),
        "t3ns_davidson_refined": (C.c_int, [P, P, c_dbl_p, c_dbl_p, C.c_int, C.c_int, C.c_double,
                                            C.c_double, C.c_int, c_dbl_p, c_int_p, c_int_p]),
        "t3ns_davidson_dev": (C.c_int, [C.c_int64, P, P, DEV_MATVEC, P, P, c_dbl_p, C.c_int,
                                        C.c_int, C.c_double, C.c_int, c_int_p]),
        "t3ns_heff_use_graphs": (C.c_int, [P, C.c_int]),
        "t3ns_davidson_dev_noalloc": (C.c_int, [C.c_int64, P, P, DEV_MATVEC, P, P, c_dbl_p,
                                                C.c_int, C.c_int, C.c_double, C.c_int, c_int_p]),
        "t3ns_heff_size": (C.c_int64, [P]),
        "t3ns_heff_stats": (C.c_int, [P, c_i64_p, c_i64_p, c_i64_p, c_dbl_p]),
        "t3ns_heff_plan_table": (C.c_int64, [P, c_int_p, c_dbl_p]),
        "t3ns_heff_kernel_stats": (C.c_int, [P, c_dbl_p, C.c_int]),
        "t3ns_heff_autotune": (C.c_int, [P, C.c_int, c_dbl_p, C.c_int]),
        "t3ns_heff_batch_create": (C.c_int, [C.POINTER(P)]),
        "t3ns_heff_batch_destroy": (C.c_int, [P]),
        "t3ns_heff_batch_matvec": (C.c_int, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.c_int,
                                             P]),
        "t3ns_heff_batch_matvec_host": (C.c_int, [P, C.POINTER(P), C.POINTER(c_dbl_p),
                                                  C.POINTER(c_dbl_p), C.c_int]),
        "t3ns_bupdate_plan_stats": (C.c_int, [C.POINTER(BUpdateDesc), C.POINTER(RopsDesc), c_dbl_p,
                                              C.c_int]),
        "t3ns_ckpt_write": (C.c_int, [P, C.c_char_p, C.c_int, c_int_p, C.POINTER(c_i64_p),
                                      C.POINTER(c_i64_p), C.POINTER(c_dbl_p)]),
        "t3ns_ckpt_open": (C.c_int, [C.POINTER(P), C.c_char_p]),
        "t3ns_ckpt_info": (C.c_int, [P, c_int_p, c_int_p, c_int_p, c_int_p]),
        "t3ns_ckpt_site": (C.c_int, [P, C.c_int, c_int_p, c_int_p, C.POINTER(c_i64_p),
                                     C.POINTER(c_i64_p), C.POINTER(c_dbl_p)]),
        "t3ns_ckpt_context": (C.c_int, [P, C.POINTER(P)]),
        "t3ns_ckpt_close": (C.c_int, [P]),
        "t3ns_heff_get_tuning": (C.c_int, [P, c_dbl_p]),
        "t3ns_heff_set_tuning": (C.c_int, [P, c_dbl_p]),
        "t3ns_tune_cache_clear": (None, []),
        "t3ns_plan_table": (C.c_int64, [C.POINTER(HeffDesc), C.POINTER(RopsDesc),
                                        C.POINTER(SiteDesc), c_int_p, c_dbl_p, c_dbl_p]),
        "t3ns_plan_block_dims": (C.c_int, [C.POINTER(HeffDesc), C.POINTER(RopsDesc),
                                           C.POINTER(SiteDesc), c_int_p]),
        "t3ns_ctx_create": (C.c_int, [C.POINTER(P), C.c_char_p, C.c_char_p, c_dbl_p, C.c_int,
                                      c_int_p, c_int_p, C.c_int, C.c_int, C.c_int]),
        "t3ns_ctx_create_qc": (C.c_int, [C.POINTER(P), C.c_char_p, C.c_char_p, c_int_p, c_int_p,
                                         C.c_int, C.c_int, C.c_int]),
        "t3ns_ctx_sweep_steps": (C.c_int, [P, C.c_int, c_int_p, c_int_p, C.c_int]),
        "t3ns_ctx_destroy": (C.c_int, [P]),
        "t3ns_ctx_info": (C.c_int, [P, c_int_p, c_int_p, c_int_p, c_int_p, c_int_p]),
        "t3ns_ctx_bond": (C.c_int, [P, C.c_int, c_int_p, c_int_p, c_int_p, C.c_int]),
        "t3ns_ctx_bond_sites": (C.c_int, [P, C.c_int, c_int_p]),
        "t3ns_ctx_set_bond": (C.c_int, [P, C.c_int, C.c_int, c_int_p, c_int_p]),
        "t3ns_ctx_fill_seeded": (C.c_int, [P, C.c_int]),
        "t3ns_ctx_site_tensor": (C.c_int, [P, C.c_int, C.c_uint64, c_int_p, C.POINTER(c_i64_p),
                                           C.POINTER(c_i64_p), C.POINTER(c_dbl_p)]),
        "t3ns_ctx_site_bonds": (C.c_int, [P, C.c_int, c_int_p, c_int_p]),
        "t3ns_ctx_append_desc": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, C.POINTER(AppendDesc),
                                           C.POINTER(RopsDesc)]),
        "t3ns_ctx_pupdate_desc": (C.c_int, [P, C.c_int, C.c_int, C.c_uint64, C.POINTER(PUpdateDesc)]),
        "t3ns_step_random": (C.c_int, [C.POINTER(P), P, c_int_p, C.c_int, C.c_uint64]),
        "t3ns_step_destroy": (C.c_int, [P]),
        "t3ns_step_size": (C.c_int64, [P]),
        "t3ns_step_heff_desc": (C.c_int, [P, C.POINTER(HeffDesc), C.POINTER(RopsDesc),
                                          C.POINTER(SiteDesc)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().t3ns_last_error().decode(errors="replace")
        raise RuntimeError("%s failed (%s): %s" % (what, ERRORS.get(rc, rc), msg))
    return rc


def exported_symbols():
    """Names of the C-ABI entry points declared in include/*.h."""
    import re
    names = []
    inc = os.path.join(os.path.dirname(_HERE), "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(t3ns_[a-z0-9_]+)\s*\(", txt, re.M)
    return sorted(set(names))


def split_view_dict(v):
    """The arrays of a split result view as numpy copies."""
    def arr(p, n, dt):
        return np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].astype(dt).copy() if n else np.zeros(0, dt)
    nU = int(v.U_beginblock[v.U_nrblocks]) if v.U_nrblocks else 0
    nV = int(v.V_beginblock[v.V_nrblocks]) if v.V_nrblocks else 0
    ns = arr(v.nsing, v.nrsectors, np.int32)
    return {
        "U_nrsites": v.U_nrsites,
        "U_qnumbers": arr(v.U_qnumbers, v.U_nrblocks * v.U_nrsites, np.int64),
        "U_beginblock": arr(v.U_beginblock, v.U_nrblocks + 1, np.int64),
        "U_tel": arr(v.U_tel, nU, np.float64),
        "V_qnumbers": arr(v.V_qnumbers, v.V_nrblocks, np.int64),
        "V_beginblock": arr(v.V_beginblock, v.V_nrblocks + 1, np.int64),
        "V_tel": arr(v.V_tel, nV, np.float64),
        "bond_irreps": arr(v.bond_irreps, v.bond_nrSecs * MAX_SYMMETRIES, np.int32).reshape(-1, MAX_SYMMETRIES),
        "bond_dims": arr(v.bond_dims, v.bond_nrSecs, np.int32),
        "nsing": ns,
        "kept": arr(v.kept, v.nrsectors, np.int32),
        "sing": arr(v.sing, int(ns.sum()), np.float64),
        "norm": (v.norm[0], v.norm[1]),
        "entropy": (v.entropy[0], v.entropy[1]),
        "norm_U": v.norm_U,
        "sweeps": v.sweeps,
    }
