"""The C driver t3ns_amd/t3ns_optimize_step (driver/optimize_step.c) replays
the reference's optimize_siteTensor call sequence (src/optimize_network.c:
188-226) through the C ABI itself: t3ns_heff_init, t3ns_make_diagonal,
t3ns_sparse_eigensolve with t3ns_matvecT3NS as the host callback,
t3ns_heff_destroy, free; then the device-callback solve (t3ns_davidson_dev
with t3ns_heff_dev_matvec).  On the GPU its energies and eigenvector must
equal those of the ctypes path (Heff.davidson on the same step and guess)
to 1e-12 relative (energy) and 1e-10 (vector, up to sign), with the same
status and iteration count."""
import json
import os
import subprocess

import numpy as np
import pytest

import t3ns_amd as T
from conftest import GOLDEN, NETW, ROOT, netw_text

DRIVER = os.path.join(ROOT, "t3ns_amd", "t3ns_optimize_step")


def test_driver_is_built_and_checks_its_arguments():
    assert os.access(DRIVER, os.X_OK), "run t3ns_amd/build.py"
    r = subprocess.run([DRIVER], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


CASES = [
    # (network path, model args, symmetries, target, D, sites)
    ("networks/10_T3NS", ["--hubbard", "1,4"], ["U1", "U1"], [5, 5], 24, [0, 1, 10, 2]),
    ("networks/10_T3NS", ["--hubbard", "1,4"], ["U1", "SU2"], [10, 0], 24, [1, 10]),
    ("ref_networks/10_T3NS", ["--fcidump", os.path.join(GOLDEN, "fcidumps", "N2.STO3G.FCIDUMP")],
     ["U1", "SU2"], [14, 0], 64, [0, 1, 2, 3]),
]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-%s" % (c[0].split("/")[-1], "x".join(c[2])))
def test_driver_matches_ctypes_path(gpu, case, tmp_path):
    netw, model, sy, ts, D, sites = case
    print("driver case", case[0], sy, sites, flush=True)
    path = os.path.join(GOLDEN, netw + ".netw")
    if model[0] == "--fcidump":
        with open(model[1]) as f:
            ctx = T.Context(netw_text(netw), "QC", [], sy, ts, D, fcidump=f.read())
    else:
        t, U = map(float, model[1].split(","))
        ctx = T.Context(netw_text(netw), "NN_HUBBARD", [t, U], sy, ts, D)
    st = T.Step(ctx, sites, seed=7)
    h = T.Heff.from_step(st)
    x0 = np.random.default_rng(5).standard_normal(h.size)
    guess, out = tmp_path / "guess.bin", tmp_path / "out.bin"
    x0.tofile(guess)
    cmd = [DRIVER, "--netw", path] + model + ["--sym", ",".join(sy), "--ts",
                                              ",".join(map(str, ts)), "--D", str(D), "--sites",
                                              ",".join(map(str, sites)), "--seed", "7", "--guess",
                                              str(guess), "--tol", "1e-7", "--its", "500",
                                              "--out", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["size"] == h.size
    e, x, its, rc = h.davidson(x0, 30, 2, 1e-7, 500)
    # same iteration on the same operator: the same status (0 converged, 1
    # at the iteration limit; the seeded operators are not a Hamiltonian, so
    # convergence is not the point)
    assert res["status"] == rc and res["status_dev"] == rc and res["its_dev"] == its
    assert abs(res["energy"] - e) <= 1e-12 * abs(e), (res["energy"], e)
    assert abs(res["energy_dev"] - e) <= 1e-12 * abs(e), (res["energy_dev"], e)
    xv = np.fromfile(out, dtype=np.float64)
    s = np.sign(np.dot(xv, x))
    assert np.max(np.abs(xv - s * x)) < 1e-10
