"""Synthetic FCIDUMP Hamiltonians for the quantum-chemistry workloads.

One-electron h is symmetric with h_ii ~ U[-2.0, -0.5] and h_ij ~ N(0, 0.1^2);
the two-electron integrals are a low-rank positive factorisation
(ij|kl) = sum_q B_qij B_qkl with B_q = (G_q + G_q^T) / 2, G entries ~ N(0, s^2),
which carries the 8-fold permutational symmetry of real orbitals by
construction.  With a point group, orbitals get irreps round-robin and every
integral whose irrep product is not totally symmetric is zeroed (XOR of the
psi4-ordered labels, the abelian product used in src/symmetry_pg.c).

The text follows the FCIDUMP layout read at src/hamiltonian_qc.c:783-900:
header with NORB/NELEC/MS2/ORBSYM/ISYM, terminator line, then
"value i j k l" lines in chemists' notation (1-based), one per 8-fold class
(i>=j, k>=l, ij>=kl), one-electron lines with k=l=0 and the core energy last.
"""
import numpy as np

# psi4 irrep index -> Molpro (FCIDUMP ORBSYM) label, inverse of the table in
# src/symmetry_pg.c:41-89 (fcidump_to_psi4) for D2h.
_PSI4_TO_MOLPRO_D2H = [1, 4, 6, 7, 8, 5, 3, 2]
_NR_IRREPS = {"C1": 1, "Ci": 2, "C2": 2, "Cs": 2, "D2": 4, "C2v": 4, "C2h": 4, "D2h": 8}


def integrals(norb, seed, rank=None, scale=0.05, pg="C1"):
    """Returns (h[norb,norb], eri[norb,norb,norb,norb] in (ij|kl) order,
    psi4 irreps of the orbitals)."""
    rank = 2 * norb if rank is None else rank
    rng = np.random.default_rng(seed)
    h = rng.normal(0.0, 0.1, size=(norb, norb))
    h = np.triu(h, 1)
    h = h + h.T
    h[np.diag_indices(norb)] = rng.uniform(-2.0, -0.5, size=norb)
    G = rng.normal(0.0, scale, size=(rank, norb, norb))
    B = 0.5 * (G + G.transpose(0, 2, 1))
    eri = np.einsum("qij,qkl->ijkl", B, B, optimize=True)
    nirr = _NR_IRREPS[pg]
    irreps = np.arange(norb) % nirr
    if nirr > 1:
        ii = irreps
        h = h * (ii[:, None] == ii[None, :])
        prod = ii[:, None, None, None] ^ ii[None, :, None, None] ^ ii[None, None, :, None] ^ \
            ii[None, None, None, :]
        eri = eri * (prod == 0)
    return h, eri, irreps


def _molpro_labels(irreps, pg):
    if pg == "C1":
        return [1] * len(irreps)
    if pg != "D2h":
        raise ValueError("only C1 and D2h labels are generated")
    return [_PSI4_TO_MOLPRO_D2H[i] for i in irreps]


def fcidump_text(h, eri, nelec, irreps=None, pg="C1", core=0.0, ms2=0, tol=0.0):
    """FCIDUMP text for (h, eri); eri is (ij|kl) with 8-fold symmetry."""
    n = h.shape[0]
    if irreps is None:
        irreps = np.zeros(n, dtype=int)
    labels = _molpro_labels(irreps, pg)
    out = [" &FCI NORB= %d,NELEC= %d,MS2= %d," % (n, nelec, ms2),
           "  ORBSYM=" + ",".join(str(x) for x in labels) + ",",
           "  ISYM=1,", " /"]
    # canonical 8-fold representatives: i>=j, k>=l, (i,j)>=(k,l)
    i, j = np.tril_indices(n)
    pair = np.stack([i, j], 1)                   # pairs with i>=j, compound index order
    P = len(pair)
    a, b = np.tril_indices(P)                    # pair a >= pair b
    I, J = pair[a, 0], pair[a, 1]
    K, L = pair[b, 0], pair[b, 1]
    v = eri[I, J, K, L]
    keep = np.abs(v) > tol
    lines = ["%.16E %4d %4d %4d %4d" % t for t in zip(v[keep], I[keep] + 1, J[keep] + 1,
                                                         K[keep] + 1, L[keep] + 1)]
    hv = h[i, j]
    keep = np.abs(hv) > tol
    lines += ["%.16E %4d %4d %4d %4d" % (x, p + 1, q + 1, 0, 0)
              for x, p, q in zip(hv[keep], i[keep], j[keep])]
    lines.append("%.16E %4d %4d %4d %4d" % (core, 0, 0, 0, 0))
    return "\n".join(out + lines) + "\n"


def qc_workload(norb, nelec, seed, rank=None, pg="C1"):
    """FCIDUMP text of the synthetic Hamiltonian used by the QC workloads."""
    h, eri, irr = integrals(norb, seed, rank=rank, pg=pg)
    return fcidump_text(h, eri, nelec, irr, pg=pg)
