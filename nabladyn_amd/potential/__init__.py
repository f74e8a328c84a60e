"""Potentials: models whose one scalar energy defines forces and virial (LJ, tabulated and
sums of them)."""

from ._base import Potential
from ._pair import LennardJones, Tabulated, lennard_jones, pair_index, tabulated
from ._sum import Sum, sum  # noqa: A004 (the reference's name)

__all__ = ("Potential", "LennardJones", "Tabulated", "Sum", "lennard_jones", "tabulated", "sum",
           "pair_index")
