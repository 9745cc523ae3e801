"""BF16 matvec (T3NS_BF16: operator blocks stored in BF16, each divided by
a power-of-two scale at or above its largest magnitude, the scales folded
into the term prefactors; FP32 matrix-core arithmetic and accumulation)
against the oracle (FP64).  Tolerance: relative 1e-2 in the 2-norm (BF16
keeps 8 significant bits, so each operator value carries up to 2^-9
relative rounding, three operators per term).  The error must also exceed
the FP32 path's (the operators really are BF16), and the BF16 operator
serves as the low-precision stage of the refined Davidson, whose energy is
an FP64 one (within 1e-6 Eh of plain FP64)."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from conftest import netw_text
from test_qc_heff import fcidump

TOL_BF16 = 1e-2


def _hub(case, seed=5):
    name, sy, ts, D, sites = case
    ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    return ctx, T.Step(ctx, sites, seed=seed)


@pytest.mark.gpu
@pytest.mark.parametrize("case", HEFF_CASES, ids=case_id)
def test_matvec_bf16_hubbard(gpu, case):
    ctx, st = _hub(case)
    print("bf16 case", case_id(case), flush=True)
    d, o, s = st.descriptors()
    oh = O.OracleHeff(d, o, s)
    x = np.random.default_rng(3).standard_normal(oh.size)
    yr = oh.matvec(x)
    errs = {}
    for prec in (T.BF16, T.FP32):
        h = T.Heff(d, o, s, prec, keepalive=st)
        errs[prec] = np.linalg.norm(h.matvec(x) - yr) / np.linalg.norm(yr)
        h.close()
    assert errs[T.BF16] < TOL_BF16, errs
    assert errs[T.BF16] > 10 * errs[T.FP32], errs


@pytest.mark.gpu
@pytest.mark.parametrize("case", QC_CASES, ids=case_id)
def test_matvec_bf16_qc(gpu, case):
    name, sy, ts, D, sites = case
    print("bf16 qc case", case_id(case), flush=True)
    ctx = T.Context(netw_text(name), "QC", [], sy, ts, D, fcidump=fcidump())
    st = T.Step(ctx, sites, seed=7)
    d, o, s = st.descriptors()
    h = T.Heff(d, o, s, T.BF16, keepalive=st)
    oh = O.OracleHeff(d, o, s)
    x = np.random.default_rng(2).standard_normal(h.size)
    yr = oh.matvec(x)
    assert np.linalg.norm(h.matvec(x) - yr) / np.linalg.norm(yr) < TOL_BF16


@pytest.mark.gpu
@pytest.mark.parametrize("placement", ["lds", "ldsw", "global"])
def test_matvec_bf16_placements(gpu, placement, monkeypatch):
    """Every kernel path with BF16 operators."""
    monkeypatch.setenv("T3NS_HEFF_PLACEMENT", placement)
    ctx, st = _hub(HEFF_CASES[4])
    d, o, s = st.descriptors()
    oh = O.OracleHeff(d, o, s)
    x = np.random.default_rng(17).standard_normal(oh.size)
    yr = oh.matvec(x)
    h = T.Heff(d, o, s, T.BF16, keepalive=st)
    if placement != "lds":
        assert h.kernel_stats()["pairs_global_x"] > 0
    assert np.linalg.norm(h.matvec(x) - yr) / np.linalg.norm(yr) < TOL_BF16


@pytest.mark.gpu
def test_refined_solve_with_bf16_stage(gpu):
    """Davidson on BF16 matvecs, refined with FP64 ones, on the N2/STO-3G
    step [4, 5] over isometric device environments: the energy matches
    the plain FP64 solve within 1e-6 Eh."""
    from t3ns_amd.environment import Environment
    from env_util import isometric_tensors
    ctx = T.Context(netw_text("ref_networks/10_DMRG"), "QC", [], ["U1", "SU2"], [14, 0], 1000,
                    fcidump=fcidump())
    st = T.Step(ctx, [4, 5], seed=3)
    env = Environment(ctx, seed=5, site_tensor=isometric_tensors(9))
    d, o, s = env.step_operators(st)
    h64 = T.Heff(d, o, s, T.FP64, keepalive=(st, env))
    h16 = T.Heff(d, o, s, T.BF16, keepalive=(st, env))
    guess = np.random.default_rng(3).standard_normal(h64.size)
    e64, _, its, rc = h64.davidson(guess, tol=1e-8, max_its=400)
    assert rc == 0
    eb, _, its_lo, its_hi, rc = T.davidson_refined(h16, h64, guess, tol=1e-8, switch_tol=1e-4,
                                                   max_its=400)
    print("E64 %.12f bf16-refined %.12f its %d / %d+%d" % (e64, eb, its, its_lo, its_hi))
    assert rc == 0
    assert abs(eb - e64) < 1e-6
