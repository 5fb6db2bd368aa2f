"""Reference-format input decks written from the seeded workloads, for the
deck reader's tests.  The species carry names and element compositions
(H2/O2/H2O/OH/H/O/HO2/H2O2/N2) so that the reader derives their molecular
weights; their NASA-7 and Lennard-Jones data are the synthetic sets of
combustioncodegpu_amd.cases."""
import numpy as np

from combustioncodegpu_amd import cases, deck

NAMES = ["H2", "O2", "H2O", "OH", "H", "O", "HO2", "H2O2", "N2"]
ELEMENTS = ["O", "H", "N"]
COMP = [[0, 2, 0], [2, 0, 0], [1, 2, 0], [1, 1, 0], [0, 1, 0], [1, 0, 0], [2, 1, 0], [2, 2, 0], [0, 0, 2]]


def hot_record(ns=9):
    y = np.zeros(ns)
    y[[1, 2, 8]] = [0.2, 0.1, 0.7]
    return dict(mach=-1.0, velocity=80.0, pressure=1.2 * 101325.0, temperature=1800.0, u=0.6, v=0.8, w=0.0, yk=y)


def walls_deck(root, mx=20, my=16, mz=12, box=True):
    """c1_walls geometry and conditions; initialised with the inflow record,
    plus (box=True) a hotter group in x in [4, 10] mm, y in [3, 9] mm."""
    case = cases.c1_walls(mx, my, mz)
    init = [(case.inflow[0], None)]
    if box:
        init.append((hot_record(), ((0.004, 0.010), (0.003, 0.009), (-1.0, 1.0))))
    deck.write_deck(root, case, species_names=NAMES, elements=ELEMENTS, elem_comp=COMP, init=init)
    return case, init


def lattice_deck(root, sizes=((9, 8), (7,), (6,))):
    """Two blocks joined by inner faces, outflow outside, one group."""
    case = cases.lattice([list(s) for s in sizes])
    init = [(hot_record(), None)]
    deck.write_deck(root, case, species_names=NAMES, elements=ELEMENTS, elem_comp=COMP, init=init)
    return case, init
