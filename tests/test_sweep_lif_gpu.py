"""The reference's LiF regression energies (tests/test3.c.in, test4.c.in)
from device two-site sweeps on its lif_T3NS network (25 orbitals, 6
electrons, Z2 x U1 x SU2 x C2v) under the reference's own optimisation
scheme (tests/test3.c.in:81-84, test4.c.in:81-84; struct regime,
include/optScheme.h:28-45): regime 1 at minD = maxD = 100, truncation 1e-4
'E', Davidson tolerance 1e-6, 4 Davidson iterations per step, at most 2
sweeps, E_CONV 1e-4; regime 2 at D = 140 (test4: 100), 100 iterations, at
most 10 sweeps, E_CONV 1e-6.  Energies within the reference's tolerance
1e-5.  The start state is a seeded canonical one (the reference's is
random), so the path to the minimum differs, not the scheme."""
import gzip
import os

import pytest

import t3ns_amd as T
from conftest import GOLDEN
from t3ns_amd.sweep import DeviceSolver, Regime, Sweeper

CASES = [
    # (target (N, 2S, C2v irrep), D of the second regime, energy, source)
    ([6, 0, 0], 140, -107.1152169273, "tests/test3.c.in:86"),
    ([6, 2, 0], 140, -106.8754093145, "tests/test3.c.in:86"),
    ([6, 0, 1], 100, -106.8889308194337247, "tests/test4.c.in:86"),
]


def lif_scheme(D2):
    """The two regimes of tests/test3.c.in:81-84 (second regime D2)."""
    return [Regime(100, 100, 1e-4, 1e-6, davidson_max_its=4, max_sweeps=2, e_conv=1e-4),
            Regime(D2, D2, 1e-4, 1e-6, davidson_max_its=100, max_sweeps=10, e_conv=1e-6)]


def _inputs():
    with gzip.open(os.path.join(GOLDEN, "fcidumps", "LiF_3.05.FCIDUMP.gz"), "rt") as f:
        fci = f.read()
    with open(os.path.join(GOLDEN, "ref_networks", "lif_T3NS.netw")) as f:
        netw = f.read()
    return netw, fci


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "ts%s-D%d" % ("_".join(map(str, c[0])), c[1]))
def test_device_sweep_lif_reference_energy(gpu, case):
    target, D2, want, src = case
    print("case lif", target, D2, flush=True)
    netw, fci = _inputs()
    ctx = T.Context(netw, "QC", [], ["U1", "SU2", "C2v"], target, 100, fcidump=fci)
    dm = Sweeper(ctx, 100, seed=1, solver=DeviceSolver("f64"))
    e = dm.run_scheme(lif_scheme(D2))
    print("sweep energies", ["%.10f" % v for v in dm.energies], flush=True)
    print("E %.10f reference %.10f (%s)" % (e, want, src))
    assert abs(e - want) < 1e-5
