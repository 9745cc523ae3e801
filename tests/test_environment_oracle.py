"""Operators of the Hamiltonian itself, built bond by bond as init_rops does
(src/optimize_network.c:428-468) through t3ns_amd.environment with the
oracle's renormalisation (append_phys, update_physical, update_branching)
and the oracle's Heff.

What the mathematics fixes: the effective Hamiltonian of any step is
symmetric (the environment is T^dagger H T for whatever tensors the updates
used), and with isometric tensors over untruncated sectors it is the
Hamiltonian on the step's whole space, so its lowest eigenvalue is the
exact ground-state energy (Hubbard, by exact diagonalisation here) or, with
truncated sectors, an upper bound to it (N2/STO-3G, whose full-CI energy
-107.648250974014 the reference's own tests/test1.c.in:99 holds).  A wrong
sign, prefactor, index or coupling symbol anywhere in the chain breaks one
of these."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle as O
import t3ns_amd as T
from conftest import netw_text
from env_util import (ORACLE_BACKEND, dense, hubbard_edges, hubbard_ground_energy,
                      isometric_tensors)
from t3ns_amd.environment import Environment
from test_qc_heff import fcidump

# One branching site (6) joining the chains 0-1, 2-3 and 4-5.
TREE7 = """NR_SITES = 7
NR_PHYS_SITES = 6
NR_BONDS = 9
&END
0 1 2 3 4 5 *
&END
-1\t0
0\t1
1\t6
-1\t2
2\t3
3\t6
6\t4
4\t5
5\t-1
"""

N2_FCI = -107.648250974014  # tests/test1.c.in:99


def network(name):
    return TREE7 if name == "tree7" else netw_text(name)


def test_ed_two_sites_closed_form():
    """Two sites, one electron of each spin: E0 = (U - sqrt(U^2 + 16 t^2)) / 2."""
    for t, U in ((1.0, 0.0), (1.0, 4.0), (0.5, 8.0)):
        want = (U - np.sqrt(U * U + 16 * t * t)) / 2
        assert abs(hubbard_ground_energy(2, [(0, 1)], 1, 1, t, U) - want) < 1e-12


@pytest.mark.parametrize("edges", [[(i, i + 1) for i in range(5)],
                                   [(0, 1), (2, 3), (4, 5), (1, 3), (1, 4), (3, 4)]],
                         ids=["chain", "tree"])
def test_ed_free_fermions(edges):
    """U = 0: the sum of the lowest one-particle levels of -t A per spin
    (fails without the Jordan-Wigner signs on the tree's long hops)."""
    A = np.zeros((6, 6))
    for i, j in edges:
        A[i, j] = A[j, i] = 1.0
    eps = np.linalg.eigvalsh(-A)
    for nu, nd in ((3, 3), (2, 4), (4, 1)):
        want = eps[:nu].sum() + eps[:nd].sum()
        assert abs(hubbard_ground_energy(6, edges, nu, nd, 1.0, 0.0) - want) < 1e-12


def test_tree_hopping_graph():
    ctx = T.Context(TREE7, "NN_HUBBARD", [1.0, 4.0], ["U1", "U1"], [3, 3], 64)
    assert hubbard_edges(ctx) == [(0, 1), (1, 3), (1, 4), (2, 3), (3, 4), (4, 5)]


HUB_STEPS = [("6_DMRG", [2, 3]), ("tree7", [1, 3, 6, 4]), ("tree7", [1, 6, 4]), ("tree7", [3, 6]),
             ("tree7", [6, 4])]
HUB_SYMS = [(["U1", "U1"], [3, 3]), (["U1", "SU2"], [6, 0])]


@pytest.mark.parametrize("sy", HUB_SYMS, ids=lambda s: "x".join(s[0]))
@pytest.mark.parametrize("net", HUB_STEPS, ids=lambda n: "%s-%s" % (n[0], "_".join(map(str, n[1]))))
def test_hubbard_step_energy_is_exact(net, sy):
    name, sites = net
    syms, target = sy
    ctx = T.Context(network(name), "NN_HUBBARD", [1.0, 4.0], syms, target, 256)
    st = T.Step(ctx, sites, seed=3)
    env = Environment(ctx, seed=5, site_tensor=isometric_tensors(9), **ORACLE_BACKEND)
    d, o, s = env.step_operators(st)
    H = dense(O.OracleHeff(d, o, s).matvec, st.size)
    assert np.linalg.norm(H - H.T) <= 1e-13 * np.linalg.norm(H)
    e0 = np.linalg.eigvalsh(H)[0]
    exact = hubbard_ground_energy(6, hubbard_edges(ctx), 3, 3, 1.0, 4.0)
    assert abs(e0 - exact) < 1e-10


@pytest.mark.parametrize("sy", [(["U1", "SU2", "D2h"], [14, 0, 0]), (["U1", "SU2"], [14, 0])],
                         ids=lambda s: "x".join(s[0]))
def test_qc_step_energy_bounds_full_ci(sy):
    """N2/STO-3G on the reference's 10_DMRG network, step [4, 5], isometric
    environments (the large side truncated to its sector dims): a
    variational energy, at most 0.05 Eh above the full-CI energy."""
    syms, target = sy
    ctx = T.Context(netw_text("ref_networks/10_DMRG"), "QC", [], syms, target, 1000,
                    fcidump=fcidump())
    st = T.Step(ctx, [4, 5], seed=3)
    env = Environment(ctx, seed=5, site_tensor=isometric_tensors(9), **ORACLE_BACKEND)
    d, o, s = env.step_operators(st)
    oh = O.OracleHeff(d, o, s)
    n = st.size
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    hx, hy = oh.matvec(x), oh.matvec(y)
    assert abs(x @ hy - y @ hx) <= 1e-13 * np.linalg.norm(x) * np.linalg.norm(hy)
    A = sla.LinearOperator((n, n), matvec=oh.matvec, dtype=float)
    e0 = sla.eigsh(A, k=1, which="SA", tol=1e-12)[0][0]
    assert N2_FCI - 1e-9 <= e0 < N2_FCI + 0.05


SYM_CASES = [
    ("hub", "10_T3NS", ["U1", "SU2"], [10, 0], 24, [0, 1, 10, 2]),
    ("hub", "10_T3NS", ["U1", "U1"], [5, 5], 24, [10, 2]),
    ("hub", "16_T3NS", ["U1", "SU2"], [16, 0], 32, [3, 4]),
    ("qc", "ref_networks/10_T3NS", ["U1", "SU2", "D2h"], [14, 0, 0], 48, [5, 6, 7]),
    ("qc", "ref_networks/10_T3NS", ["U1", "U1"], [7, 7], 32, [2, 3]),
]


@pytest.mark.parametrize("case", SYM_CASES, ids=lambda c: "%s-%s-%s-%s" % (
    c[0], c[1].split("/")[-1], "x".join(c[2]), "_".join(map(str, c[5]))))
def test_step_heff_symmetric_with_branching_updates(case):
    """Seeded (non-isometric) tensors on tree networks, so the chain runs
    update_branching too: x.Hy == y.Hx."""
    kind, name, syms, target, D, sites = case
    if kind == "hub":
        ctx = T.Context(netw_text(name), "NN_HUBBARD", [1.0, 4.0], syms, target, D)
    else:
        ctx = T.Context(netw_text(name), "QC", [], syms, target, D, fcidump=fcidump())
    st = T.Step(ctx, sites, seed=3)
    env = Environment(ctx, seed=5, **ORACLE_BACKEND)
    d, o, s = env.step_operators(st)
    oh = O.OracleHeff(d, o, s)
    rng = np.random.default_rng(2)
    for _ in range(2):
        x, y = rng.standard_normal(st.size), rng.standard_normal(st.size)
        hx, hy = oh.matvec(x), oh.matvec(y)
        assert abs(x @ hy - y @ hx) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(hy)
