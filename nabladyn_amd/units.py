"""Units and physical constants (nm, ps, Da, kJ/mol, K, e), as the reference's ``units.py``."""

AVOGADRO = 6.02214076e23  # 1/mol
BOLTZMANN = 0.00831446261815324  # kJ/(mol K)
COULOMB = 138.93545764438198  # kJ nm/(mol e^2)
ANGSTROM = 0.1  # nm
FS = 0.001  # ps
KCAL_PER_MOL = 4.184  # kJ/mol
EV = 96.48533212331002  # kJ/mol
BAR = 0.0602214076  # kJ/(mol nm^3)
ATM = 1.01325 * BAR
G_PER_CM3 = 602.214076  # Da/nm^3

__all__ = ("AVOGADRO", "BOLTZMANN", "COULOMB", "ANGSTROM", "FS", "KCAL_PER_MOL", "EV", "BAR",
           "ATM", "G_PER_CM3")
