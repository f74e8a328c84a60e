"""The sum of models (``potential/_sum/_potential.py``, built by ``potential/_api.py:273-293``):
one energy, the sum of the parts'.

A part reads the shortest given list that reaches its own cutoff (``list_for``); pairs beyond a
part's cutoff count nothing for it, so a list per cutoff changes the cost, never the result.
A sum of fused models is evaluated by ``quantity`` and the MD step part by part with the fused
kernels; any other part makes the sum an ordinary differentiable energy.
"""

import torch
from torch import Tensor, nn

from ..neighbors import NeighborPairs
from ._base import Potential


def _signature(model) -> tuple:
    return (model.num_atoms, model.dtype, model.device)


def _require_matching(models) -> None:
    """The parts share atoms, dtype and device (``potential/_sum/_potential.py:10-21``)."""
    first = models[0]
    for model in models[1:]:
        if _signature(model) != _signature(first):
            raise ValueError(
                f"the models of a sum must match: {first.num_atoms} atoms in {first.dtype} on "
                f"{first.device}, against {model.num_atoms} in {model.dtype} on "
                f"{model.device}; convert a sum as a whole, not one of its models",
            )


def list_for(pairs: NeighborPairs, cutoff: float) -> NeighborPairs:
    """Of the outer list and its inner lists, the one with the smallest cutoff at least
    ``cutoff``, carrying the shorter lists as its own inner lists
    (``potential/_sum/_potential.py:34-43``). The outer list always qualifies."""
    candidates = sorted((*pairs.inner, pairs), key=lambda listed: listed.cutoff)
    chosen = next(listed for listed in candidates if listed.cutoff >= cutoff)
    inner = tuple(listed for listed in candidates if listed.cutoff < chosen.cutoff)
    return NeighborPairs(chosen.index, chosen.num_atoms, chosen.cutoff, inner, chosen.full)


# the former name, kept for callers
own_pairs = list_for


class Sum(Potential):
    """Models whose energies add, E = sum_k E_k, so F and W add too (``potential.sum``)."""

    def __init__(self, *models: Potential):
        super().__init__()
        if any(not isinstance(model, Potential) for model in models):
            raise TypeError("every model of a sum must be a Potential")
        if len(models) < 2:
            raise ValueError(f"a sum needs at least two models, not {len(models)}")
        _require_matching(models)
        self.parts = nn.ModuleList(models)
        self._cutoff = max(model.cutoff for model in models)

    @property
    def cutoff(self) -> float:
        return self._cutoff

    @property
    def num_atoms(self) -> int:
        return self.parts[0].num_atoms

    # a part converted on its own after the sum was built would leave the parts mismatched;
    # dtype and device are read only through this check
    @property
    def dtype(self) -> torch.dtype:
        _require_matching(self.parts)
        return self.parts[0].dtype

    @property
    def device(self) -> torch.device:
        _require_matching(self.parts)
        return self.parts[0].device

    @property
    def _list_cutoffs(self) -> tuple[float, ...]:
        return tuple(sorted({c for part in self.parts for c in part._list_cutoffs}))

    @property
    def _fused(self) -> bool:
        return all(part._fused for part in self.parts)

    def _native(self):
        """The C-ABI description the MD step takes for a sum of two LJ or tabulated models that
        share their atom types: the first model's ``ndn_potential`` with ``next`` at the
        second's. Any other sum raises ``TypeError``."""
        import ctypes

        from ._pair import _PairModel

        first, second = (self.parts[0], self.parts[1]) if len(self.parts) == 2 else (None, None)
        if not (isinstance(first, _PairModel) and isinstance(second, _PairModel)):
            raise TypeError("the MD step evaluates a sum of exactly two LennardJones or "
                            "Tabulated models")
        if first._num_types != second._num_types or not torch.equal(first.atom_types,
                                                                    second.atom_types):
            raise TypeError("the MD step evaluates a sum of two models with the same atom types")
        head, keep_head = first._native()
        tail, keep_tail = second._native()
        # equal types: the first model's array serves both
        tail.atom_types = head.atom_types
        head.next = ctypes.cast(ctypes.pointer(tail), ctypes.c_void_p)
        return head, (keep_head, keep_tail, tail)

    @property
    def atom_types(self) -> Tensor:
        """The parts' atom types (those of the first)."""
        return self.parts[0].atom_types

    @property
    def _num_types(self) -> int:
        return getattr(self.parts[0], "_num_types", 1)

    def _trained_masks(self) -> dict[str, Tensor]:
        masks = {}
        for k, part in enumerate(self.parts):
            masks.update({f"parts.{k}.{name}": mask for name, mask in part._trained_masks().items()})
        return masks

    def _energy_wrapped(self, positions: Tensor, cell: Tensor, pairs: NeighborPairs,
                        strain: Tensor | None = None) -> Tensor:
        total = None
        for part in self.parts:
            e = part._energy_wrapped(positions, cell, list_for(pairs, part.cutoff), strain)
            total = e if total is None else total + e
        return total


def sum(*models: Potential) -> Potential:  # noqa: A001 (the reference's name)
    """A model whose energy is the sum of the models' energies (``potential/_api.py:273-293``).

    E = sum_k E_k over two or more ``models``, so its forces and virial are the sums of theirs;
    its cutoff is the largest of theirs. Each model reads the shortest given list that reaches
    its cutoff (``NeighborList(..., inner_cutoffs=...)`` gives one per cutoff), so a
    short-range model beside a long-range one costs about what it costs alone; a list without
    inner lists serves every model at the cost of the longest. The models must share their
    atoms, dtype and device; the sum holds them, not copies (parameters ``parts.<k>.<name>``).
    No host sync.

    Raises ``TypeError`` for a model that is not a Potential, and ``ValueError`` for fewer than
    two models, or for models of other atom counts, dtypes or devices, when the sum is built
    or, after one model was converted alone, used.
    """
    return Sum(*models)


__all__ = ("Sum", "list_for", "own_pairs", "sum")
