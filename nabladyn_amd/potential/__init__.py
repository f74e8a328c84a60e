"""Potentials: models whose one scalar energy defines forces and virial (LJ and tabulated)."""

from ._base import Potential
from ._pair import LennardJones, Tabulated, lennard_jones, pair_index, tabulated

__all__ = ("Potential", "LennardJones", "Tabulated", "lennard_jones", "tabulated", "pair_index")
