"""GPU parity of the Heff matvec and diagonal against the oracle (CPU
restatement of src/Heff.c) on the same inputs.  FP64: relative error 1e-12
in the 2-norm and per element 1e-12 of the largest element; FP32: 2e-5."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, case_id
from conftest import netw_text
from plan_util import without_zero_couplings


def _setup(case, seed=5):
    name, sy, ts, D, sites = case
    ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], sy, ts, D)
    st = T.Step(ctx, sites, seed=seed)
    d, o, s = st.descriptors()
    return ctx, st, d, o, s


@pytest.mark.gpu
@pytest.mark.parametrize("case", HEFF_CASES, ids=case_id)
def test_matvec_fp64_matches_oracle(gpu, case):
    ctx, st, d, o, s = _setup(case)
    print("case", case_id(case), flush=True)
    h = T.Heff(d, o, s, T.FP64, keepalive=st)
    oh = O.OracleHeff(d, o, s)
    rng = np.random.default_rng(11)
    for _ in range(2):
        x = rng.standard_normal(h.size)
        y = h.matvec(x)
        yr = oh.matvec(x)
        err = np.linalg.norm(y - yr) / np.linalg.norm(yr)
        assert err < 1e-12, err
        assert np.max(np.abs(y - yr)) <= 1e-12 * np.max(np.abs(yr))
    # same plan table on the device object
    ints, prefs = h.plan_table()
    oi, op, _ = without_zero_couplings(*oh.plan_table())
    assert np.array_equal(ints, oi)


@pytest.mark.gpu
@pytest.mark.parametrize("case", HEFF_CASES, ids=case_id)
def test_diagonal_matches_oracle(gpu, case):
    ctx, st, d, o, s = _setup(case)
    h = T.Heff(d, o, s, T.FP64, keepalive=st)
    oh = O.OracleHeff(d, o, s)
    dg = h.diagonal()
    dr = oh.diagonal()
    np.testing.assert_allclose(dg, dr, rtol=1e-12, atol=1e-12 * np.max(np.abs(dr)))


@pytest.mark.gpu
@pytest.mark.parametrize("case", HEFF_CASES[:5], ids=case_id)
def test_matvec_fp32(gpu, case):
    ctx, st, d, o, s = _setup(case)
    h = T.Heff(d, o, s, T.FP32, keepalive=st)
    oh = O.OracleHeff(d, o, s)
    x = np.random.default_rng(3).standard_normal(h.size)
    y = h.matvec(x)
    yr = oh.matvec(x)
    assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 2e-5


@pytest.mark.gpu
@pytest.mark.parametrize("placement", ["lds", "ldsw", "global"])
@pytest.mark.parametrize("case", HEFF_CASES, ids=case_id)
def test_matvec_placements(gpu, case, placement, monkeypatch):
    """Every kernel path forced on every pair: input blocks staged in LDS
    where they fit (lds), read from global memory (ldsw, the default), and
    stage outputs in global scratch as well (global)."""
    monkeypatch.setenv("T3NS_HEFF_PLACEMENT", placement)
    ctx, st, d, o, s = _setup(case)
    print("case", case_id(case), placement, flush=True)
    oh = O.OracleHeff(d, o, s)
    x = np.random.default_rng(17).standard_normal(oh.size)
    yr = oh.matvec(x)
    for prec, tol in ((T.FP64, 1e-12), (T.FP32, 2e-5)):
        h = T.Heff(d, o, s, prec, keepalive=st)
        st_ = h.kernel_stats()
        if placement != "lds":
            assert st_["pairs_global_x"] > 0
        if placement == "global":
            assert st_["pairs_global_w"] > 0
        y = h.matvec(x)
        assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < tol
        h.close()


@pytest.mark.gpu
def test_placement_option_rejects_unknown(gpu, monkeypatch):
    monkeypatch.setenv("T3NS_HEFF_PLACEMENT", "bogus")
    ctx, st, d, o, s = _setup(HEFF_CASES[0])
    with pytest.raises(RuntimeError):
        T.Heff(d, o, s, T.FP64, keepalive=st)


@pytest.mark.gpu
@pytest.mark.parametrize("case", HEFF_CASES, ids=case_id)
def test_matvec_reference_order(gpu, case, monkeypatch):
    """The reference's contraction orders (T3NS_HEFF_ORDER=reference) give
    the same matvec as the cost-model orders (default), both against the
    oracle (FP64 1e-12)."""
    ctx, st, d, o, s = _setup(case)
    oh = O.OracleHeff(d, o, s)
    x = np.random.default_rng(23).standard_normal(oh.size)
    yr = oh.matvec(x)
    for order in ("reference", "model"):
        monkeypatch.setenv("T3NS_HEFF_ORDER", order)
        h = T.Heff(d, o, s, T.FP64, keepalive=st)
        y = h.matvec(x)
        assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-12, order
        h.close()
