"""Two-site DMRG sweeps on the device (t3ns_amd.sweep.Sweeper with the
device makesiteTensor, renormalisation, Heff + Davidson and split): the
reference's N2/STO-3G regression energy -107.648250974014 within its
tolerance 1e-8 (tests/test1.c.in:99), FP64 and FP32 matvecs with FP64
refinement, on the reference's 10_T3NS tree (its test1 network) and 10-site
chain with the four symmetry sets of tests/test1.c.in:34-41; and the
10-site Hubbard chain at D = 64 against the same sweep run on the oracle
(1e-8 Eh)."""
import pytest

import t3ns_amd as T
from conftest import netw_text
from env_util import OracleSolver
from t3ns_amd.sweep import Sweeper, DeviceSolver
from test_environment_oracle import N2_FCI, TREE7
from env_util import hubbard_edges, hubbard_ground_energy
from test_qc_heff import fcidump


# the four symmetry sets of tests/test1.c.in:34-41
TEST1_SYMS = [(["U1", "U1"], [7, 7]), (["U1", "U1", "D2h"], [7, 7, 0]), (["U1", "SU2"], [14, 0]),
              (["U1", "SU2", "D2h"], [14, 0, 0])]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "mixed"])
@pytest.mark.parametrize("sy", TEST1_SYMS, ids=lambda s: "x".join(s[0]))
@pytest.mark.parametrize("net", ["10_T3NS", "10_DMRG"])
def test_device_sweep_n2_reference_energy(gpu, net, sy, prec):
    syms, target = sy
    print("case n2", net, syms, prec, flush=True)
    ctx = T.Context(netw_text("ref_networks/" + net), "QC", [], syms, target, 1000,
                    fcidump=fcidump())
    dm = Sweeper(ctx, 1000, seed=1, solver=DeviceSolver(prec, mixed_min_size=0))
    e = dm.run(4, e_conv=1e-10)
    print("E %.12f sweeps %d" % (e, len(dm.energies)))
    assert abs(e - N2_FCI) < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_device_sweep_hubbard_10_matches_oracle(gpu, prec):
    print("case hubbard 10 D=64", prec, flush=True)

    def sweep(solver):
        ctx = T.Context(netw_text("10_DMRG"), "NN_HUBBARD", [1.0, 4.0], ["U1", "U1"], [5, 5], 64)
        dm = Sweeper(ctx, 64, seed=1, solver=solver)
        return dm.run(3)
    e_dev = sweep(DeviceSolver(prec, mixed_min_size=0))
    e_ref = sweep(OracleSolver())
    print("E device %.12f oracle %.12f" % (e_dev, e_ref))
    assert abs(e_dev - e_ref) < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("site_size", [3, 4])
@pytest.mark.parametrize("sy", [TEST1_SYMS[0], TEST1_SYMS[3]], ids=lambda s: "x".join(s[0]))
def test_device_multisite_sweep_n2(gpu, sy, site_size):
    """Three- and four-site steps around the branching sites of the
    reference's 10_T3NS network (HOSVD by repeated splits)."""
    syms, target = sy
    print("case n2 multisite", syms, site_size, flush=True)
    ctx = T.Context(netw_text("ref_networks/10_T3NS"), "QC", [], syms, target, 1000,
                    fcidump=fcidump())
    dm = Sweeper(ctx, 1000, seed=1, solver=DeviceSolver("f64"), site_size=site_size)
    e = dm.run(4, e_conv=1e-10)
    print("E %.12f sweeps %d" % (e, len(dm.energies)))
    assert abs(e - N2_FCI) < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_device_sweep_hubbard_tree_exact(gpu, prec):
    """NN_HUBBARD (t = 1, U = 4) on a 7-site tree with one branching site,
    U1 x SU2, untruncated: the device sweep ends on the exact energy of the
    tree's hopping graph (1e-9)."""
    print("case hubbard tree7", prec, flush=True)
    ctx = T.Context(TREE7, "NN_HUBBARD", [1.0, 4.0], ["U1", "SU2"], [6, 0], 256)
    dm = Sweeper(ctx, 256, seed=1, solver=DeviceSolver(prec, mixed_min_size=0))
    e = dm.run(3, e_conv=1e-12)
    exact = hubbard_ground_energy(6, hubbard_edges(ctx), 3, 3, 1.0, 4.0)
    print("E %.12f exact %.12f" % (e, exact))
    assert abs(e - exact) < 1e-9

