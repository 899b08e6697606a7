"""Element masses used by the workloads, as liblpmd/support/elements.cc:20-35 lists them
(the masses BasicAtom takes from ElemMass[Z], liblpmd/lpmd/basicatom.h:57-67).
Test infrastructure only."""
import numpy as np

MASS = {
    1: 1.00794,      # H
    11: 22.989768,   # Na
    13: 26.981539,   # Al
    17: 35.4527,     # Cl
    18: 39.948,      # Ar
    26: 55.847,      # Fe
    28: 58.6934,     # Ni
    29: 63.546,      # Cu
    42: 95.94,       # Mo
    47: 107.8682,    # Ag
    79: 196.96654,   # Au
}


def mass_of(z):
    z = np.atleast_1d(np.asarray(z))
    return np.array([MASS[int(q)] for q in z], dtype=np.float64)
