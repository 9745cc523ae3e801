"""The synthetic FCIDUMP generator of the QC workloads (t3ns_amd.synthetic):
8-fold permutational symmetry of the two-electron integrals, the stated
distributions' ranges, D2h masking (an integral survives only when the XOR
of its orbitals' psi4 irreps is the totally symmetric irrep; one-electron
terms only within an irrep), and the text round trip of the canonical
8-fold representatives."""
import numpy as np

from t3ns_amd import synthetic


def test_eri_symmetry_and_ranges():
    h, eri, irr = synthetic.integrals(8, seed=1, rank=16)
    assert np.allclose(h, h.T)
    assert np.all((np.diag(h) >= -2.0) & (np.diag(h) <= -0.5))
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.allclose(eri, eri.transpose(perm))
    # positive semidefinite as a (ij),(kl) matrix: a sum of squares
    w = np.linalg.eigvalsh(eri.reshape(64, 64))
    assert w.min() > -1e-12


def test_d2h_masking():
    h, eri, irr = synthetic.integrals(12, seed=2, pg="D2h")
    assert list(irr[:9]) == [0, 1, 2, 3, 4, 5, 6, 7, 0]
    i = np.arange(12)
    assert np.all(h[irr[:, None] != irr[None, :]] == 0)
    prod = irr[:, None, None, None] ^ irr[None, :, None, None] ^ irr[None, None, :, None] ^ \
        irr[None, None, None, :]
    assert np.all(eri[prod != 0] == 0)
    assert np.count_nonzero(eri[prod == 0]) > 0
    assert np.allclose(eri, eri.transpose(2, 3, 0, 1))
    del i


def test_fcidump_text_round_trip():
    h, eri, irr = synthetic.integrals(5, seed=3, rank=4)
    txt = synthetic.fcidump_text(h, eri, 4, irr)
    lines = txt.splitlines()
    assert lines[0].startswith(" &FCI NORB= 5,NELEC= 4")
    body = [ln.split() for ln in lines[4:]]
    h2 = np.zeros_like(h)
    e2 = np.zeros_like(eri)
    for v, p, q, r, s in body:
        p, q, r, s, v = int(p), int(q), int(r), int(s), float(v)
        if p == 0:
            continue
        if r == 0:
            h2[p - 1, q - 1] = h2[q - 1, p - 1] = v
            continue
        for a, b, c, d in [(p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                           (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)]:
            e2[a - 1, b - 1, c - 1, d - 1] = v
    assert np.allclose(h2, h, atol=1e-15)
    assert np.allclose(e2, eri, atol=1e-15)
