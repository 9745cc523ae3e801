"""Two-site DMRG sweeps (t3ns_amd.sweep.Sweeper, the optimisation loop of
src/optimize_network.c:488-570) run entirely on the oracle: makesiteTensor,
operator renormalisation, the Heff eigenpair and the split with truncation.

What the sweep must give: the exact ground-state energy where nothing is
truncated (the 6-site Hubbard chain against exact diagonalisation), a
variational energy that does not rise from sweep to sweep where it is
(10 sites at D = 16), and the reference's own regression energy for
N2/STO-3G, -107.648250974014 within its tolerance 1e-8
(tests/test1.c.in:99; D = 1000 keeps every state), on its 10-site chain
and on its 10_T3NS tree."""
import pytest

import t3ns_amd as T
from conftest import netw_text
from env_util import OracleSolver, hubbard_edges, hubbard_ground_energy
from t3ns_amd.sweep import Sweeper
from test_environment_oracle import N2_FCI, TREE7
from test_qc_heff import fcidump


@pytest.mark.parametrize("sy", [(["U1", "U1"], [3, 3]), (["U1", "SU2"], [6, 0])],
                         ids=lambda s: "x".join(s[0]))
def test_sweep_hubbard_6_exact(sy):
    syms, target = sy
    ctx = T.Context(netw_text("6_DMRG"), "NN_HUBBARD", [1.0, 4.0], syms, target, 64)
    dm = Sweeper(ctx, 64, seed=1, solver=OracleSolver())
    e = dm.run(2)
    exact = hubbard_ground_energy(6, hubbard_edges(ctx), 3, 3, 1.0, 4.0)
    assert abs(e - exact) < 1e-10


def test_sweep_hubbard_10_truncated_is_variational():
    ctx = T.Context(netw_text("10_DMRG"), "NN_HUBBARD", [1.0, 4.0], ["U1", "U1"], [5, 5], 16)
    dm = Sweeper(ctx, 16, seed=1, solver=OracleSolver())
    dm.run(3)
    exact = -5.380618820415  # sparse ED of the 10-site chain (tests/test_environment_gpu.py)
    es = dm.energies
    assert all(b <= a + 1e-10 for a, b in zip(es, es[1:]))
    assert exact - 1e-10 <= es[-1] < exact + 0.01


@pytest.mark.parametrize("case", [("10_DMRG", ["U1", "SU2", "D2h"], [14, 0, 0]),
                                  ("10_T3NS", ["U1", "SU2"], [14, 0])],
                         ids=lambda c: "%s-%s" % (c[0], "x".join(c[1])))
def test_sweep_n2_reproduces_reference_energy(case):
    """On the reference's 10_T3NS network the sweep passes its branching
    sites (steps of a physical and a branching site, branching updates)."""
    name, syms, target = case
    ctx = T.Context(netw_text("ref_networks/" + name), "QC", [], syms, target, 1000,
                    fcidump=fcidump())
    dm = Sweeper(ctx, 1000, seed=1, solver=OracleSolver())
    e = dm.run(4, e_conv=1e-10)
    assert abs(e - N2_FCI) < 1e-8


@pytest.mark.parametrize("site_size", [3, 4])
def test_sweep_multisite_steps_n2(site_size):
    """Three- and four-site steps around the branching sites of 10_T3NS,
    decomposed by repeated splits (HOSVD, src/siteTensor_decompose.c:
    1533-1585)."""
    ctx = T.Context(netw_text("ref_networks/10_T3NS"), "QC", [], ["U1", "SU2"], [14, 0], 1000,
                    fcidump=fcidump())
    dm = Sweeper(ctx, 1000, seed=1, solver=OracleSolver(), site_size=site_size)
    assert any(len(st) == site_size for st in dm.steps)
    e = dm.run(3, e_conv=1e-10)
    assert abs(e - N2_FCI) < 1e-8


@pytest.mark.parametrize("sy", [(["U1", "U1"], [3, 3]), (["U1", "SU2"], [6, 0])],
                         ids=lambda s: "x".join(s[0]))
def test_sweep_hubbard_tree_exact(sy):
    """NN_HUBBARD on a 7-site tree with one branching site (two-site steps,
    the sweep passes it with branching updates): nothing is truncated, so
    the sweep ends on the exact energy of the tree's hopping graph.  (Its
    three-site sweep order holds a {branching, physical, physical} chain,
    which the multisite structure does not take; the makenetwork.py trees
    give {physical, branching, physical} objects only.)"""
    syms, target = sy
    ctx = T.Context(TREE7, "NN_HUBBARD", [1.0, 4.0], syms, target, 256)
    dm = Sweeper(ctx, 256, seed=1, solver=OracleSolver())
    e = dm.run(3, e_conv=1e-12)
    exact = hubbard_ground_energy(6, hubbard_edges(ctx), 3, 3, 1.0, 4.0)
    assert abs(e - exact) < 1e-10


def test_scheme_n2_two_regimes():
    """Two regimes as in tests/test1.c.in:86-89 (a first truncated one, then
    one keeping every state): the lowest energy is the full-CI energy."""
    ctx = T.Context(netw_text("ref_networks/10_T3NS"), "QC", [], ["U1", "SU2"], [14, 0], 64,
                    fcidump=fcidump())
    dm = Sweeper(ctx, 64, seed=1, solver=OracleSolver())
    e = dm.run_scheme([(16, 16, 1e-4, 1e-6, 2, 2, 1e-6), (1000, 1000, 1e-4, 1e-8, 4, 10, 1e-10)])
    assert dm.energies[0] > N2_FCI + 1e-6  # the first regime truncates
    assert abs(e - N2_FCI) < 1e-8

