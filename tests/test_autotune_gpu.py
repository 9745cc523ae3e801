"""Autotuner of the Heff work lists (t3ns_heff_autotune): whatever setting
it picks, and every setting it can pick, the matvec stays within the FP64
parity bound of the oracle (1e-12 relative); a second object with the same
block-size histogram takes the recorded setting without timing."""
import numpy as np
import pytest

import oracle as O
import t3ns_amd as T
from cases import QC_CASES, case_id
from test_qc_heff import _setup

# three-site objects: middle legs of a few states (the fused form applies)
CASES = [QC_CASES[4], QC_CASES[8], QC_CASES[1]]


def _check(h, oh, seed):
    x = np.random.default_rng(seed).standard_normal(h.size)
    y, yr = h.matvec(x), oh.matvec(x)
    err = np.linalg.norm(y - yr) / np.linalg.norm(yr)
    assert err < 1e-12, err
    assert np.max(np.abs(y - yr)) <= 1e-12 * np.max(np.abs(yr))


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=case_id)
def test_autotune_keeps_parity_and_records(gpu, case):
    T.clear_tune_cache()
    ctx, st = _setup(case)
    print("case", case_id(case), flush=True)
    d, o, s = st.descriptors()
    h = T.Heff(d, o, s, T.FP64, keepalive=st)
    oh = O.OracleHeff(d, o, s)
    r = h.autotune(max_trials=6)
    assert 1 <= r["trials"] <= 6 and not r["cached"]
    assert r["ms"] > 0 and r["ms"] <= r["baseline_ms"] * 1.02
    assert h.tuning() == {k: r[k] for k in T.Heff.TUNING}
    _check(h, oh, 1)
    # the same structure again: recorded setting, no timing
    h2 = T.Heff(d, o, s, T.FP64, keepalive=st)
    r2 = h2.autotune(max_trials=6)
    assert r2["cached"] and r2["trials"] == 0
    assert {k: r2[k] for k in T.Heff.TUNING} == {k: r[k] for k in T.Heff.TUNING}
    _check(h2, oh, 2)


@pytest.mark.gpu
@pytest.mark.parametrize("setting", [
    dict(chunk_steps=300, fuse_over=1.0, fuse=1),
    dict(chunk_steps=3000, fuse_over=4.0, fuse=1),
    dict(chunk_steps=1000, fuse_over=0.0, fuse=1),
    dict(chunk_steps=1000, fuse_over=2.0, fuse=0),
    dict(chunk_steps=1000, fuse_over=2.0, fuse=1, sched=2),
], ids=lambda t: "-".join("%s%g" % kv for kv in t.items()))
def test_every_setting_keeps_parity(gpu, setting):
    ctx, st = _setup(QC_CASES[8])
    d, o, s = st.descriptors()
    oh = O.OracleHeff(d, o, s)
    for prec, tol in ((T.FP64, 1e-12), (T.FP32, 2e-5)):
        h = T.Heff(d, o, s, prec, keepalive=st)
        h.set_tuning(**setting)
        assert h.tuning()["fuse"] == setting["fuse"]
        x = np.random.default_rng(4).standard_normal(h.size)
        y, yr = h.matvec(x), oh.matvec(x)
        assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < tol
