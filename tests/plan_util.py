"""Plan-table comparison helper.

The reference drops a plan entry when its symmetry prefactor is below 1e-20
(src/Heff.c:827).  A coupling that is exactly zero comes out of the Wigner
9j evaluation as rounding noise (~1e-18, its sign and size depending on the
evaluation order), so whether the reference keeps it is arbitrary; the
product drops every prefactor below 1e-14.  The oracle follows the
reference's rule, so its table is compared after removing those entries."""
import numpy as np

ZERO_COUPLING = 1e-14


def without_zero_couplings(ints, prefs):
    """(rows, prefactors, number removed) of an oracle plan table."""
    keep = np.abs(prefs) >= ZERO_COUPLING
    return ints[keep], prefs[keep], int(np.count_nonzero(~keep))
