"""Pins of the oracle's symmetry layer (CPU).

* Wigner 6j/9j of the oracle against sympy's exact symbols (the reference's
  src/Wigner.c computes the same Racah sums in long double).
* Every prefactor the Heff/renormalisation path uses (src/symmetries.c:158-332)
  against the reference's own symmetry sources compiled unchanged into
  oracle/_ref/libref_sym.so, on random irreps of Z2 x U1 x SU2 (x D2h) and
  Z2 x U1 x U1.
A wrong sign, a transposed irrep index or a dropped factor in the oracle
fails these on some sample."""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle as O

Z2, U1, SU2, D2h = 0, 1, 2, 10


def _sympy():
    try:
        from sympy.physics.wigner import wigner_6j, wigner_9j
        from sympy import Rational
        return wigner_6j, wigner_9j, Rational
    except ImportError:  # pragma: no cover
        pytest.skip("sympy not available")


def test_wigner6j_matches_sympy():
    w6, _, R = _sympy()
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(3000):
        j = rng.integers(0, 7, size=6)
        try:
            exact = float(w6(*[R(int(v), 2) for v in j]))
        except ValueError:  # sympy rejects non-triangular sets: the symbol is 0
            exact = 0.0
        got = O.lib().oracle_wigner6j(*[int(v) for v in j])
        assert abs(got - exact) < 1e-13, (j, got, exact)
        checked += exact != 0
    assert checked > 20


def test_wigner9j_matches_sympy():
    _, w9, R = _sympy()
    rng = np.random.default_rng(1)
    checked = 0
    for _ in range(1500):
        j = rng.integers(0, 4, size=9)
        try:
            exact = float(w9(*[R(int(v), 2) for v in j], prec=None))
        except ValueError:
            exact = 0.0
        got = O.lib().oracle_wigner9j(*[int(v) for v in j])
        assert abs(got - exact) < 1e-13, (j, got, exact)
        checked += exact != 0
    assert checked > 5


def _ref():
    r = O.ref()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return r


IP = C.POINTER(C.c_int)


def _irreps(rng, sgs, n):
    """n random irreps as rows of MAX_SYMMETRIES ints."""
    out = np.zeros((n, 5), dtype=np.int32)
    for k, g in enumerate(sgs):
        if g == Z2:
            out[:, k] = rng.integers(0, 2, size=n)
        elif g == U1:
            out[:, k] = rng.integers(-4, 5, size=n)
        elif g == SU2:
            out[:, k] = rng.integers(0, 7, size=n)
        else:
            out[:, k] = rng.integers(0, 8, size=n)
    return out


def _tri(rng, a, b):
    """A spin coupling a and b (doubled values): |a-b|, |a-b|+2, ..., a+b."""
    return int(rng.choice(np.arange(abs(a - b), a + b + 1, 2)))


def _coupled_su2(rng, irr, col):
    """Overwrite column col of a 3x3 irrep table with SU2 values whose rows
    and columns all satisfy the triangle rule (a nonzero 9j is possible)."""
    for _ in range(100):
        a, b, d, e = (int(v) for v in rng.integers(0, 5, size=4))
        c, f = _tri(rng, a, b), _tri(rng, d, e)
        g, h = _tri(rng, a, d), _tri(rng, b, e)
        lo, hi = max(abs(c - f), abs(g - h)), min(c + f, g + h)
        ks = [k for k in range(lo, hi + 1) if (k - c - f) % 2 == 0 and (k - g - h) % 2 == 0]
        if ks:
            k = int(rng.choice(ks))
            vals = [a, b, c, d, e, f, g, h, k]
            for i in range(9):
                irr[i, col] = vals[i]
            return irr
    return irr


def _ptrs(rows):
    arr = (IP * len(rows))()
    for i, r in enumerate(rows):
        arr[i] = r.ctypes.data_as(IP)
    return arr


SYMSETS = [[Z2, U1, SU2], [Z2, U1, U1], [Z2, U1, SU2, D2h]]


@pytest.mark.parametrize("sgs", SYMSETS, ids=["U1xSU2", "U1xU1", "U1xSU2xD2h"])
def test_prefactors_match_reference(sgs):
    ref = _ref()
    L = O.lib()
    rng = np.random.default_rng(len(sgs) * 7 + sgs[2])
    csgs = (C.c_int * len(sgs))(*sgs)
    ns = len(sgs)
    fns = {
        "pAppend": (9, [0, 1], lambda irr, p: L.oracle_prefactor_pAppend(irr.ctypes.data_as(IP), p, csgs, ns),
                    lambda rows, p: ref.prefactor_pAppend(_ptrs(rows), p, csgs, ns)),
        "pUpdate": (9, [0, 1], lambda irr, p: L.oracle_prefactor_pUpdate(irr.ctypes.data_as(IP), p, csgs, ns),
                    lambda rows, p: ref.prefactor_pUpdate(_ptrs(rows), p, csgs, ns)),
        "bUpdate": (9, [0, 1, 2], lambda irr, p: L.oracle_prefactor_bUpdate(irr.ctypes.data_as(IP), p, csgs, ns),
                    lambda rows, p: ref.prefactor_bUpdate(_ptrs(rows), p, csgs, ns)),
        "adjoint": (3, [ord(c) for c in "c123"],
                    lambda irr, p: L.oracle_prefactor_adjoint(irr.ctypes.data_as(IP), C.c_char(bytes([p])), csgs, ns),
                    lambda rows, p: ref.prefactor_adjoint(_ptrs(rows), C.c_char(bytes([p])), csgs, ns)),
        "mirror": (3, [0], lambda irr, p: L.oracle_prefactor_mirror_coupling(irr.ctypes.data_as(IP), csgs, ns),
                   lambda rows, p: ref.prefactor_mirror_coupling(_ptrs(rows), csgs, ns)),
        "add_P": (6, [0, 1], lambda irr, p: L.oracle_prefactor_add_P_operator(irr.ctypes.data_as(IP), p, csgs, ns),
                  lambda rows, p: ref.prefactor_add_P_operator(_ptrs(rows), p, csgs, ns)),
    }
    for name, (nirr, params, fo, fr) in fns.items():
        getattr(ref, {"pAppend": "prefactor_pAppend", "pUpdate": "prefactor_pUpdate",
                      "bUpdate": "prefactor_bUpdate", "adjoint": "prefactor_adjoint",
                      "mirror": "prefactor_mirror_coupling",
                      "add_P": "prefactor_add_P_operator"}[name]).restype = C.c_double
        nonzero = 0
        for trial in range(300):
            irr = _irreps(rng, sgs, nirr)
            if nirr == 9 and SU2 in sgs and trial % 2:
                irr = _coupled_su2(rng, irr, sgs.index(SU2))
            rows = [irr[i] for i in range(nirr)]
            for p in params:
                a = fo(irr, p)
                b = fr(rows, p)
                assert abs(a - b) <= 1e-13 * max(1.0, abs(b)), (name, p, irr[:, :ns].tolist(), a, b)
                nonzero += abs(b) > 0
        assert nonzero > 0, name
    # combine_MPOs: [2][3] irreps and 3 MPO irreps
    ref.prefactor_combine_MPOs.restype = C.c_double
    for trial in range(300):
        irr = _irreps(rng, sgs, 6)
        mpo = _irreps(rng, sgs, 3)
        for isdmrg, extra in itertools.product([0, 1], [0, 2]):
            a = L.oracle_prefactor_combine_MPOs(irr.ctypes.data_as(IP), mpo.ctypes.data_as(IP),
                                                csgs, ns, isdmrg, extra)
            b = ref.prefactor_combine_MPOs(_ptrs([irr[i] for i in range(6)]),
                                           _ptrs([mpo[i] for i in range(3)]), csgs, ns, isdmrg,
                                           extra)
            assert abs(a - b) <= 1e-13 * max(1.0, abs(b)), ("combine", isdmrg, extra, a, b)
