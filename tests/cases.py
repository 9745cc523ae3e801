"""Shared parity cases: (network, symmetries, target, max_dim, sites)."""
HEFF_CASES = [
    ("8_DMRG", ["U1", "U1"], [4, 4], 16, [3, 4]),
    ("8_DMRG", ["U1", "SU2"], [8, 0], 24, [3, 4]),
    ("10_DMRG", ["U1", "U1"], [5, 5], 64, [4, 5]),
    ("10_T3NS", ["U1", "U1"], [5, 5], 24, [0, 10, 2]),
    ("10_T3NS", ["U1", "U1"], [5, 5], 24, [0, 1, 10, 2]),
    ("10_T3NS", ["U1", "U1"], [5, 5], 24, [10, 2]),
    ("10_T3NS", ["U1", "SU2"], [10, 0], 24, [0, 1, 10, 2]),
    ("10_T3NS", ["U1", "SU2"], [10, 0], 24, [1, 10]),
    ("16_T3NS", ["U1", "SU2"], [16, 0], 48, [3, 4]),
]


def case_id(c):
    return "%s-%s-D%d-%s" % (c[0], "x".join(c[1]), c[3], "_".join(map(str, c[4])))


# Quantum-chemistry cases on the reference's own test inputs
# (tests/fcidumps/N2.STO3G.FCIDUMP with tests/networks/10_T3NS.netw and
# 10_DMRG.netw, symmetry sets of tests/test1.c.in:34-41).
QC_FCIDUMP = "fcidumps/N2.STO3G.FCIDUMP"
QC_CASES = [
    ("ref_networks/10_T3NS", ["U1", "U1"], [7, 7], 64, [0, 1, 2, 3]),
    ("ref_networks/10_T3NS", ["U1", "SU2"], [14, 0], 64, [0, 1, 2, 3]),
    ("ref_networks/10_T3NS", ["U1", "SU2", "D2h"], [14, 0, 0], 64, [2, 3]),
    ("ref_networks/10_T3NS", ["U1", "U1", "D2h"], [7, 7, 0], 64, [1, 2]),
    ("ref_networks/10_T3NS", ["U1", "SU2", "D2h"], [14, 0, 0], 64, [5, 6, 7]),
    ("ref_networks/10_DMRG", ["U1", "SU2", "D2h"], [14, 0, 0], 64, [4, 5]),
    ("ref_networks/10_DMRG", ["U1", "U1"], [7, 7], 48, [2, 3]),
    # physical site on the branching site's middle leg with its first leg
    # external (the transposed/batched makesiteTensor products)
    ("ref_networks/10_T3NS", ["U1", "U1"], [7, 7], 64, [7, 8]),
    ("ref_networks/10_T3NS", ["U1", "SU2"], [14, 0], 64, [7, 8, 9]),
]
