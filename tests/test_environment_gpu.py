"""The Hamiltonian's own operators built on the device (t3ns_amd.environment
over the device append/update entry points, init_rops of
src/optimize_network.c:428-468) against the same chain on the oracle, and
ground-state energies from the device Davidson (FP64, and FP32 matvecs with
FP64 refinement) against exact diagonalisation, the reference's N2/STO-3G
full-CI bound and the oracle's own eigenvalue.

Tolerances: chained operators within 1e-11 relative of the oracle's (each
update is within 1e-12, the chain compounds a few of them); FP64 energies
within 1e-9 Eh; FP32-with-refinement energies within 1e-6 Eh of the FP64
ones."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle as O
import t3ns_amd as T
from conftest import netw_text
from env_util import ORACLE_BACKEND, hubbard_edges, hubbard_ground_energy, isometric_tensors
from t3ns_amd.environment import Environment
from test_environment_oracle import N2_FCI, SYM_CASES, network
from test_qc_heff import fcidump


def _ctx(kind, name, syms, target, D):
    if kind == "hub":
        return T.Context(network(name), "NN_HUBBARD", [1.0, 4.0], syms, target, D)
    return T.Context(netw_text(name), "QC", [], syms, target, D, fcidump=fcidump())


def _blocks(desc):
    out = []
    for k in range(desc.nrops):
        h = desc.hss_of_ops[k]
        nb = desc.begin_blocks_of_hss[h + 1] - desc.begin_blocks_of_hss[h]
        bb = np.ctypeslib.as_array(desc.beginblock[k], shape=(nb + 1,)).copy()
        n = int(bb[-1])
        tel = np.ctypeslib.as_array(desc.tel[k], shape=(max(n, 1),))[:n].copy() if n else np.zeros(0)
        out.append((bb, tel))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", SYM_CASES, ids=lambda c: "%s-%s-%s-%s" % (
    c[0], c[1].split("/")[-1], "x".join(c[2]), "_".join(map(str, c[5]))))
def test_device_environment_matches_oracle_chain(gpu, case):
    kind, name, syms, target, D, sites = case
    print("case", case, flush=True)
    ctx_d = _ctx(kind, name, syms, target, D)
    ctx_o = _ctx(kind, name, syms, target, D)
    st_d, st_o = T.Step(ctx_d, sites, seed=3), T.Step(ctx_o, sites, seed=3)
    env_d = Environment(ctx_d, seed=5)
    env_o = Environment(ctx_o, seed=5, **ORACLE_BACKEND)
    dd, od, sd = env_d.step_operators(st_d)
    do, oo, so = env_o.step_operators(st_o)
    for leg in range(3):
        assert od[leg].nrops == oo[leg].nrops and od[leg].P_operator == oo[leg].P_operator
        for (gb, gt), (wb, wt) in zip(_blocks(od[leg]), _blocks(oo[leg])):
            assert np.array_equal(gb, wb)
            if len(wt):
                assert np.max(np.abs(gt - wt)) <= 1e-11 * max(1.0, np.max(np.abs(wt)))
    h = T.Heff(dd, od, sd, T.FP64, keepalive=(st_d, env_d))
    oh = O.OracleHeff(do, oo, so)
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal(h.size), rng.standard_normal(h.size)
    hx, hy = h.matvec(x), h.matvec(y)
    assert np.linalg.norm(hx - oh.matvec(x)) <= 1e-11 * np.linalg.norm(hx)
    assert abs(x @ hy - y @ hx) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(hy)


def _energies(ctx, sites, seed=3):
    """(FP64 Davidson energy, refined FP32 energy, low/FP64 iterations) of
    the step's Heff over isometric device-built environments."""
    st = T.Step(ctx, sites, seed=seed)
    env = Environment(ctx, seed=5, site_tensor=isometric_tensors(9))
    d, o, s = env.step_operators(st)
    h64 = T.Heff(d, o, s, T.FP64, keepalive=(st, env))
    h32 = T.Heff(d, o, s, T.FP32, keepalive=(st, env))
    guess = np.random.default_rng(seed).standard_normal(h64.size)
    e64, _, its, rc = h64.davidson(guess, tol=1e-8, max_its=400)
    assert rc == 0, "FP64 Davidson did not converge in %d iterations" % its
    em, _, its_lo, its_hi, rc = T.davidson_refined(h32, h64, guess, tol=1e-8, switch_tol=1e-4,
                                                   max_its=400)
    assert rc == 0
    return e64, em, its_lo, its_hi, (st, env, h64, h32)


HUB_EXACT = [("6_DMRG", [2, 3], ["U1", "U1"], [3, 3]), ("6_DMRG", [2, 3], ["U1", "SU2"], [6, 0]),
             ("tree7", [1, 3, 6, 4], ["U1", "SU2"], [6, 0]), ("tree7", [3, 6], ["U1", "U1"], [3, 3])]


@pytest.mark.gpu
@pytest.mark.parametrize("case", HUB_EXACT, ids=lambda c: "%s-%s-%s" % (
    c[0], "_".join(map(str, c[1])), "x".join(c[2])))
def test_device_hubbard_energy_is_exact(gpu, case):
    name, sites, syms, target = case
    print("case", case, flush=True)
    ctx = T.Context(network(name), "NN_HUBBARD", [1.0, 4.0], syms, target, 256)
    exact = hubbard_ground_energy(6, hubbard_edges(ctx), 3, 3, 1.0, 4.0)
    e64, em, its_lo, its_hi, _ = _energies(ctx, sites)
    print("E64 %.12f mixed %.12f exact %.12f its %d+%d" % (e64, em, exact, its_lo, its_hi))
    assert abs(e64 - exact) < 1e-9
    assert abs(em - e64) < 1e-6


def _hubbard_chain_sparse_ed(L, n_up, n_down, t, U):
    """Lowest eigenvalue of the open Hubbard chain by sparse exact
    diagonalisation (neighbouring modes of one spin: no Jordan-Wigner sign)."""
    import itertools

    def part(n):
        states = [sum(1 << i for i in c) for c in itertools.combinations(range(L), n)]
        index = {s: k for k, s in enumerate(states)}
        rows, cols = [], []
        for k, s in enumerate(states):
            for i in range(L - 1):
                if ((s >> i) & 1) != ((s >> (i + 1)) & 1):
                    rows.append(index[s ^ (3 << i)])
                    cols.append(k)
        h = sp.csr_matrix((np.full(len(rows), -t), (rows, cols)), shape=(len(states),) * 2)
        return states, h
    su, hu = part(n_up)
    sd, hd = part(n_down)
    H = sp.kron(hu, sp.identity(len(sd))) + sp.kron(sp.identity(len(su)), hd)
    occ = np.array([bin(a & b).count("1") for a in su for b in sd], dtype=float)
    H = (H + sp.diags(U * occ)).tocsr()
    return sla.eigsh(H, k=1, which="SA", tol=1e-13)[0][0]


@pytest.mark.gpu
def test_device_hubbard_chain_10_sites(gpu):
    """The 10-site half-filled Hubbard chain (t = 1, U = 4, U1 x U1): step
    [4, 5] over untruncated environments spans the whole 63504-state sector."""
    print("case 10_DMRG hubbard exact", flush=True)
    ctx = T.Context(netw_text("10_DMRG"), "NN_HUBBARD", [1.0, 4.0], ["U1", "U1"], [5, 5], 1024)
    e64, em, its_lo, its_hi, keep = _energies(ctx, [4, 5])
    assert keep[2].size == 63504
    exact = _hubbard_chain_sparse_ed(10, 5, 5, 1.0, 4.0)
    print("E64 %.12f mixed %.12f exact %.12f its %d+%d" % (e64, em, exact, its_lo, its_hi))
    assert abs(e64 - exact) < 1e-9
    assert abs(em - e64) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("sy", [(["U1", "SU2"], [14, 0]), (["U1", "U1"], [7, 7])],
                         ids=lambda s: "x".join(s[0]))
def test_device_qc_energy(gpu, sy):
    """N2/STO-3G, step [4, 5] of the reference's 10_DMRG network: the FP64
    Davidson energy equals the oracle's lowest eigenvalue of the same
    operators (built on the oracle) and bounds the full-CI energy; the
    refined FP32 energy is within 1e-6 Eh of it."""
    syms, target = sy
    print("case qc", syms, flush=True)
    ctx = T.Context(netw_text("ref_networks/10_DMRG"), "QC", [], syms, target, 1000,
                    fcidump=fcidump())
    e64, em, its_lo, its_hi, _ = _energies(ctx, [4, 5])
    ctx_o = T.Context(netw_text("ref_networks/10_DMRG"), "QC", [], syms, target, 1000,
                      fcidump=fcidump())
    st = T.Step(ctx_o, [4, 5], seed=3)
    env = Environment(ctx_o, seed=5, site_tensor=isometric_tensors(9), **ORACLE_BACKEND)
    oh = O.OracleHeff(*env.step_operators(st))
    A = sla.LinearOperator((st.size,) * 2, matvec=oh.matvec, dtype=float)
    eo = sla.eigsh(A, k=1, which="SA", tol=1e-12)[0][0]
    print("E64 %.12f mixed %.12f oracle %.12f its %d+%d" % (e64, em, eo, its_lo, its_hi))
    assert abs(e64 - eo) < 1e-9
    assert e64 >= N2_FCI - 1e-9
    assert abs(em - e64) < 1e-6
