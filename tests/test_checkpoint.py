"""Checkpoints in the reference's HDF5 layout (T3NScalc.h5,
src/io_to_disk.c:446-487).  Pinned by (i) the file's structure against the
reference writers' groups, attributes, datasets, element types and sizes
(read back through h5dump), (ii) an exact round trip of every stored
array, and (iii) a sweep restored from a checkpoint written between sweeps
continuing exactly as the uninterrupted sweep (oracle solver, CPU)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import t3ns_amd as T
from conftest import netw_text
from env_util import OracleSolver
from t3ns_amd.checkpoint import read_checkpoint, write_checkpoint
from t3ns_amd.sweep import Sweeper
from test_qc_heff import fcidump

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def _hubbard_ctx():
    return T.Context(netw_text("10_DMRG"), "NN_HUBBARD", [1.0, 4.0], ["U1", "U1"], [5, 5], 16)


def _qc_ctx():
    return T.Context(netw_text("ref_networks/10_T3NS"), "QC", [], ["U1", "SU2", "D2h"],
                     [14, 0, 0], 64, fcidump=fcidump())


def test_round_trip_qc(tmp_path):
    ctx = _qc_ctx()
    info = ctx.info()
    rng = np.random.default_rng(3)
    tensors = {}
    from t3ns_amd.sweep import SiteData
    for s in range(info["sites"]):
        t = SiteData(*ctx.site_tensor(s, s + 1))
        t.tel[:] = rng.standard_normal(len(t.tel))
        tensors[s] = t
    path = tmp_path / "T3NScalc.h5"
    write_checkpoint(path, ctx, tensors)
    ctx2, t2, meta = read_checkpoint(path)
    assert meta["model"] == "QC" and meta["operator_groups"] == 0
    assert ctx2.info() == info
    for b in range(info["nr_bonds"]):
        assert np.array_equal(ctx2.bond_dims(b), ctx.bond_dims(b))
    assert ctx2.sweep_steps(2) == ctx.sweep_steps(2)
    for s in range(info["sites"]):
        assert np.array_equal(t2[s].qn, tensors[s].qn)
        assert np.array_equal(t2[s].bb, tensors[s].bb)
        assert np.array_equal(t2[s].tel, tensors[s].tel)
    # the restored context gives the same problems: the same Heff plan
    st1, st2 = T.Step(ctx, [2, 3], seed=4), T.Step(ctx2, [2, 3], seed=4)
    i1, p1, f1 = T.plan_table(*st1.descriptors())
    i2, p2, f2 = T.plan_table(*st2.descriptors())
    assert np.array_equal(i1, i2) and np.array_equal(p1, p2) and f1 == f2


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not installed")
def test_layout_matches_reference_writers(tmp_path):
    """Every object the reference's write_to_disk creates, with its type:
    src/io_to_disk.c:35-51 (symsecs), 87-108 (bookkeeper), 143-164
    (sparseblocks), 194-208 (siteTensor), 238-250 (T3NS), 347-357 (rOps),
    376-393 (network); src/hamiltonian.c:327-345 and
    src/hamiltonian_nn_hubbard.c:261-270 (Hamiltonian)."""
    ctx = _hubbard_ctx()
    dm = Sweeper(ctx, 16, seed=1, solver=OracleSolver())
    path = tmp_path / "T3NScalc.h5"
    dm.save(path)
    out = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True,
                         check=True).stdout
    # object path -> (kind, type, size) from the header dump
    objs, stack = {}, []
    for line in out.splitlines():
        m = re.match(r'\s*(GROUP|DATASET|ATTRIBUTE) "([^"]+)"', line)
        if m:
            kind, name = m.groups()
            depth = (len(line) - len(line.lstrip())) // 3
            stack = stack[:depth]
            stack.append(name)
            objs["/".join(stack)] = [kind, None, None]
            continue
        m = re.match(r'\s*DATATYPE\s+(\S+)', line)
        if m and stack:
            objs["/".join(stack)][1] = m.group(1)
        m = re.match(r'\s*DATASPACE\s+SIMPLE \{ \( (\d+) \)', line)
        if m and stack:
            objs["/".join(stack)][2] = int(m.group(1))
    info = ctx.info()
    nb, ns = info["nr_bonds"], info["sites"]
    I32, F64, I64 = "H5T_STD_I32LE", "H5T_IEEE_F64LE", "H5T_STD_I64LE"

    def has(path, kind, typ, size=None):
        key = "//" + path if not path.startswith("/") else "/" + path
        assert key in objs, (key, sorted(objs)[:20])
        k, t, n = objs[key]
        assert (k, t) == (kind, typ), (key, k, t)
        if size is not None:
            assert n == size, (key, n, size)

    has("/network/nr_bonds", "ATTRIBUTE", I32, 1)
    has("/network/bonds", "DATASET", I32, 2 * nb)
    has("/network/psites", "ATTRIBUTE", I32, 1)
    has("/network/sites", "ATTRIBUTE", I32, 1)
    has("/network/sitetoorb", "DATASET", I32, ns)
    has("/network/sweeplength", "ATTRIBUTE", I32, 1)
    has("/bookkeeper/nrSyms", "ATTRIBUTE", I32, 1)
    has("/bookkeeper/Max_symmetries", "ATTRIBUTE", I32, 1)
    has("/bookkeeper/sgs", "ATTRIBUTE", I32, len(info["sgs"]))
    has("/bookkeeper/target_state", "ATTRIBUTE", I32, len(info["sgs"]))
    has("/bookkeeper/nr_bonds", "ATTRIBUTE", I32, 1)
    has("/bookkeeper/psites", "ATTRIBUTE", I32, 1)
    for b in range(nb):
        n = len(ctx.bond_dims(b))
        g = "/bookkeeper/v_symsec_%d" % b
        has(g + "/nrSecs", "ATTRIBUTE", I32, 1)
        has(g + "/bond", "ATTRIBUTE", I32, 1)
        has(g + "/totaldims", "ATTRIBUTE", I32, 1)
        has(g + "/dims", "DATASET", I32, n)
        has(g + "/irreps", "DATASET", I32, 5 * n)  # MAX_SYMMETRIES = 5 (CMakeLists.txt:19)
        has(g + "/fcidims", "DATASET", F64, n)
    has("/bookkeeper/p_symsec_0/dims", "DATASET", I32)
    has("/hamiltonian/type", "ATTRIBUTE", I32, 1)
    has("/hamiltonian/hamiltonian_data/t", "ATTRIBUTE", F64, 1)
    has("/hamiltonian/hamiltonian_data/U", "ATTRIBUTE", F64, 1)
    has("/hamiltonian/hamiltonian_data/su2", "ATTRIBUTE", I32, 1)
    has("/T3NS/nrSites", "ATTRIBUTE", I32, 1)
    for s in range(ns):
        t = dm.tensors[s]
        g = "/T3NS/tensor_%d" % s
        has(g + "/nrsites", "ATTRIBUTE", I32, 1)
        has(g + "/sites", "ATTRIBUTE", I32, 1)
        has(g + "/nrblocks", "ATTRIBUTE", I32, 1)
        has(g + "/qnumbers", "DATASET", I64, t.nrblocks)
        has(g + "/block_0/nrBlocks", "ATTRIBUTE", I32, 1)
        has(g + "/block_0/beginblock", "DATASET", I32, t.nrblocks + 1)
        has(g + "/block_0/tel", "DATASET", F64, int(t.bb[-1]))
    has("/rOps/nrOps", "ATTRIBUTE", I32, 1)
    assert not any(k.startswith("//rOps/rOperator") for k in objs)
    # the scalar values
    txt = subprocess.run([H5DUMP, "-a", "/hamiltonian/type", "-a", "/network/nr_bonds",
                          "-a", "/bookkeeper/Max_symmetries", str(path)], capture_output=True,
                         text=True, check=True).stdout
    vals = [int(v) for v in re.findall(r"\(0\): (-?\d+)", txt)]
    assert vals == [2, nb, 5]  # NN_HUBBARD in enum hamtypes (include/hamiltonian.h:30)


@pytest.mark.parametrize("model", ["hubbard", "qc"])
def test_restored_sweep_continues_exactly(tmp_path, model):
    if model == "hubbard":
        make, D, name = _hubbard_ctx, 16, "hubbard"
    else:
        make, D, name = _qc_ctx, 64, "qc"
    a = Sweeper(make(), D, seed=1, solver=OracleSolver())
    a.run(2)
    b = Sweeper(make(), D, seed=1, solver=OracleSolver())
    b.sweep()
    path = tmp_path / ("%s.h5" % name)
    b.save(path)
    c = Sweeper.restore(path, D, seed=1, solver=OracleSolver())
    c.sweep()
    assert abs(c.energies[-1] - a.energies[-1]) < 1e-10, (c.energies, a.energies)


@pytest.mark.gpu
def test_restored_device_sweep_continues_exactly(gpu, tmp_path):
    """The same on the device (fixed-order sums: bitwise reproducible)."""
    print("case hubbard 10 chain D16 device sweep restore", flush=True)
    a = Sweeper(_hubbard_ctx(), 16, seed=1)
    a.run(2)
    b = Sweeper(_hubbard_ctx(), 16, seed=1)
    b.sweep()
    path = tmp_path / "dev.h5"
    b.save(path)
    c = Sweeper.restore(path, 16, seed=1)
    c.sweep()
    assert abs(c.energies[-1] - a.energies[-1]) < 1e-12, (c.energies, a.energies)
