"""The sum of models (``potential/_sum/_potential.py``): one energy, the sum of the parts'.

Each part is evaluated by the fused kernel on the shortest given list that reaches its cutoff
(the reference's ``_own_pairs``, ``potential/_sum/_potential.py:34-43``), so a list per cutoff
changes the cost, never the result: pairs beyond a part's cutoff count nothing in its kernel.
"""

import torch
from torch import Tensor, nn

from ..neighbors import NeighborPairs
from ._base import Potential


def _check_matching(models) -> None:
    """``potential/_sum/_potential.py:10-21``: the parts share atoms, dtype and device."""
    first = models[0]
    shared = (first.num_atoms, first.dtype, first.device)
    for model in models[1:]:
        if (model.num_atoms, model.dtype, model.device) != shared:
            raise ValueError(
                f"the models of a sum must match: {first.num_atoms} atoms in {first.dtype} on "
                f"{first.device}, against {model.num_atoms} in {model.dtype} on "
                f"{model.device}; convert a sum as a whole, not one of its models",
            )


def _check_models(models) -> None:
    """``potential/_sum/_potential.py:24-31``."""
    if not all(isinstance(model, Potential) for model in models):
        raise TypeError("every model of a sum must be a Potential")
    if len(models) < 2:
        raise ValueError(f"a sum needs at least two models, not {len(models)}")
    _check_matching(models)


def own_pairs(pairs: NeighborPairs, cutoff: float) -> NeighborPairs:
    """The first list, ascending, that reaches ``cutoff``, with the shorter ones inside
    (``potential/_sum/_potential.py:34-43``): ``pairs.inner`` and then ``pairs``, which
    reaches every part's cutoff."""
    lists = (*pairs.inner, pairs)
    own = [listed for listed in lists if listed.cutoff >= cutoff][0]
    shorter = tuple(listed for listed in lists if listed.cutoff < own.cutoff)
    return NeighborPairs(own.index, own.num_atoms, own.cutoff, shorter)


class Sum(Potential):
    """Models whose energies add, E = sum_k E_k, so F and W add too (``potential.sum``)."""

    def __init__(self, *models: Potential):
        super().__init__()
        _check_models(models)
        self._cutoff = max(model.cutoff for model in models)
        self.parts = nn.ModuleList(models)

    @property
    def cutoff(self) -> float:
        return self._cutoff

    @property
    def num_atoms(self) -> int:
        return self.parts[0].num_atoms

    # a part converted alone after the sum was built would mix dtypes or devices; every
    # reader of dtype and device checks first
    @property
    def dtype(self) -> torch.dtype:
        _check_matching(self.parts)
        return self.parts[0].dtype

    @property
    def device(self) -> torch.device:
        _check_matching(self.parts)
        return self.parts[0].device

    @property
    def _list_cutoffs(self) -> tuple[float, ...]:
        return tuple(sorted({value for part in self.parts for value in part._list_cutoffs}))

    def _trained_masks(self) -> dict[str, Tensor]:
        return {
            f"parts.{index}.{name}": mask
            for index, part in enumerate(self.parts)
            for name, mask in part._trained_masks().items()
        }

    def _native(self):
        raise TypeError(
            "a sum has no single kernel description: quantity evaluates its parts one by one; "
            "the MD step takes one LennardJones or Tabulated model",
        )

    def _energy_wrapped(self, positions: Tensor, cell: Tensor, pairs: NeighborPairs,
                        strain: Tensor | None = None) -> Tensor:
        energies = [part._energy_wrapped(positions, cell, own_pairs(pairs, part.cutoff), strain)
                    for part in self.parts]
        return torch.stack(energies).sum()


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


__all__ = ("Sum", "own_pairs", "sum")
