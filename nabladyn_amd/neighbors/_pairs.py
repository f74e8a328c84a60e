"""The pairs a neighbor list hands to an energy (``neighbors/_pairs.py:8-22``)."""

from typing import NamedTuple

from torch import Tensor


class FullRows(NamedTuple):
    """A list's full rows (every pair in both orientations), for the fused force kernel: row r
    is atom ``rows[r]`` (None: atom r), its k-th neighbor the atom id
    ``index[slices[r // 64] + r % 64 + 64 k]`` for k < ``counts[r]``."""

    rows: Tensor | None
    counts: Tensor
    slices: Tensor
    index: Tensor


class NeighborPairs(NamedTuple):
    """The pairs of one frame, each pair once.

    ``index`` [2, C] int32: the first columns hold each pair closer than the list radius at
    the last build, in no set orientation; the rest are self-pairs (k mod N, k mod N), which an
    energy drops by keeping only ``index[0] != index[1]``. ``num_atoms`` is N; ``cutoff`` is the
    list's cutoff in nm. ``inner`` holds the same frame's lists at smaller cutoffs, ascending.
    ``full``, an extension the reference's pairs do not have, holds the same pairs as full rows
    when the list keeps them (None otherwise); models that read ``index`` ignore it.
    """

    index: Tensor
    num_atoms: int
    cutoff: float
    inner: tuple["NeighborPairs", ...] = ()
    full: FullRows | None = None


__all__ = ("FullRows", "NeighborPairs")
