"""The neighbor list on gfx950 (``neighbors/__init__.py``: NeighborList, NeighborPairs)."""

from ._neighbor_list import NeighborList
from ._ops import ops
from ._pairs import NeighborPairs

__all__ = ("NeighborList", "NeighborPairs", "ops")
