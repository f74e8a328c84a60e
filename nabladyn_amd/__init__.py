"""nabladyn_amd: NablaDyn's neighbor-list-and-pair-force path, MI355X (gfx950) native.

Drop-in modules: ``neighbors`` (NeighborList, NeighborPairs, the ``nabladyn_neighbors`` ops),
``potential`` (lennard_jones, tabulated), ``quantity``, ``md`` (nve, nvt), ``space``,
``observables``, ``units``. Additions: ``batch`` (many systems per launch, autotuned),
``serve`` (continuous-batching request loop), ``fit`` (hand-written backward).
"""

from . import md, neighbors, observables, potential, quantity, space, units

__all__ = ("md", "neighbors", "observables", "potential", "quantity", "space", "units")
