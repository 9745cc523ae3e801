"""Derived FP32 objects (t3ns_heff_derive): an FP32 Heff made from an
FP64 one's plan and operator arenas (copied and converted on the device,
its own FP32 work lists built on the host) keeps the
FP32 parity bound of the oracle (2e-5 relative) under every grid shape and
pair placement, agrees with an FP32 object built from the descriptors, does
not depend on its base once made, and serves as the low-precision operator
of the refined Davidson (energy within 1e-8 of FP64).  Deriving from a
non-FP64 object, deriving to another precision and tuning a derived object
are refused."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import HEFF_CASES, QC_CASES, case_id
from test_heff_gpu import _setup as hub_setup
from test_qc_heff import _setup as qc_setup

CASES = [("hub", HEFF_CASES[2]), ("hub", HEFF_CASES[6]), ("qc", QC_CASES[4]), ("qc", QC_CASES[8])]


def _descriptors(kind, case):
    if kind == "hub":
        ctx, st, d, o, s = hub_setup(case)
    else:
        ctx, st = qc_setup(case)
        d, o, s = st.descriptors()
    return ctx, st, d, o, s


@pytest.mark.gpu
@pytest.mark.parametrize("placement", ["lds", "ldsw", "global"])
@pytest.mark.parametrize("wpc", [1, 2, 4])
def test_derived_fp32_parity(gpu, monkeypatch, wpc, placement):
    monkeypatch.setenv("T3NS_HEFF_PLACEMENT", placement)
    monkeypatch.setenv("T3NS_HEFF_WPC", str(wpc))
    for kind, case in CASES:
        print("derive", wpc, placement, kind, case_id(case), flush=True)
        ctx, st, d, o, s = _descriptors(kind, case)
        h64 = T.Heff(d, o, s, T.FP64, keepalive=st)
        h32 = T.Heff(d, o, s, T.FP32, keepalive=st)
        hd = h64.derive_fp32()
        assert hd.size == h64.size and hd.kernel_stats()["wpc"] == wpc
        oh = O.OracleHeff(d, o, s)
        x = np.random.default_rng(5).standard_normal(oh.size)
        yr = oh.matvec(x)
        yd = hd.matvec(x)
        assert np.linalg.norm(yd - yr) / np.linalg.norm(yr) < 2e-5
        y32 = h32.matvec(x)
        assert np.array_equal(yd, y32)  # the same lists and arenas: the same sums
        # independent of its base once made
        h64.close()
        assert np.array_equal(hd.matvec(x), yd)
        hd.close()
        h32.close()


@pytest.mark.gpu
def test_derived_fp32_in_refined_davidson(gpu):
    """N2/STO-3G step [4, 5] over isometric device environments (the
    Hamiltonian's own operators, so the operator is symmetric)."""
    from t3ns_amd.environment import Environment
    from env_util import isometric_tensors
    from conftest import netw_text
    from test_qc_heff import fcidump
    ctx = T.Context(netw_text("ref_networks/10_DMRG"), "QC", [], ["U1", "SU2"], [14, 0], 1000,
                    fcidump=fcidump())
    st = T.Step(ctx, [4, 5], seed=3)
    env = Environment(ctx, seed=5, site_tensor=isometric_tensors(9))
    d, o, s = env.step_operators(st)
    h64 = T.Heff(d, o, s, T.FP64, keepalive=(st, env))
    hd = h64.derive_fp32()
    guess = np.random.default_rng(6).standard_normal(h64.size)
    e64, _, its, rc = h64.davidson(guess, tol=1e-8, max_its=400)
    assert rc == 0
    e, _, i1, i2, rc = T.davidson_refined(hd, h64, guess, tol=1e-8, switch_tol=1e-4, max_its=400)
    assert rc == 0 and i1 > 0
    assert abs(e - e64) < 1e-8


@pytest.mark.gpu
def test_derive_refusals(gpu):
    ctx, st, d, o, s = _descriptors("hub", HEFF_CASES[2])
    h32 = T.Heff(d, o, s, T.FP32, keepalive=st)
    with pytest.raises(RuntimeError):
        h32.derive_fp32()
    h64 = T.Heff(d, o, s, T.FP64, keepalive=st)
    hd = h64.derive_fp32()
    with pytest.raises(RuntimeError):
        hd.set_tuning(wpc=1)
    with pytest.raises(RuntimeError):
        hd.autotune(max_trials=2)
