"""The Potential contract (``potential/_base.py``): a model whose one scalar energy defines
forces and virial. Here every model evaluates through the library's fused HIP kernels:
``forward`` gives E; ``nabladyn_amd.quantity`` gives E, F and W from one launch.
"""

from abc import ABC, abstractmethod

import torch
from torch import Tensor, nn

from .. import _native, space
from ..neighbors import NeighborPairs


def _check_tensor(name, value, shape, dtype, device) -> None:
    """``potential/_base.py:11-23``."""
    if not isinstance(value, Tensor) or value.shape != shape:
        raise ValueError(f"{name} must be a tensor of shape {list(shape)}")
    if value.dtype != dtype or value.device != device:
        raise ValueError(f"{name} must be {dtype} on {device} to match the model")


def _check_pairs(pairs: NeighborPairs, num_atoms: int, cutoff: float, device) -> None:
    """``potential/_base.py:26-43``."""
    for listed in (pairs, *pairs.inner):
        index = listed.index
        if index.ndim != 2 or index.shape[0] != 2 or index.device != device:
            raise ValueError(f"pairs.index must be a [2, C] tensor on {device}")
        if listed.num_atoms != num_atoms:
            raise ValueError(
                f"the pairs are for {listed.num_atoms} atoms, the model for {num_atoms}",
            )
    if pairs.cutoff < cutoff:
        raise ValueError(
            f"the neighbor list's cutoff {pairs.cutoff} nm is below the model's {cutoff} nm",
        )


class Potential(nn.Module, ABC):
    """A model of a fixed set of atoms with one scalar energy in kJ/mol.

    ``model(positions, cell, pairs)`` checks the inputs, wraps the positions into the cell and
    returns the energy; forces and virial come from ``nabladyn_amd.quantity``.
    """

    @property
    @abstractmethod
    def cutoff(self) -> float: ...

    @property
    @abstractmethod
    def num_atoms(self) -> int: ...

    @property
    @abstractmethod
    def dtype(self) -> torch.dtype: ...

    @property
    @abstractmethod
    def device(self) -> torch.device: ...

    @property
    def _list_cutoffs(self) -> tuple[float, ...]:
        return (self.cutoff,)

    @abstractmethod
    def _native(self) -> tuple[_native.Potential, tuple[Tensor, ...]]:
        """The C-ABI description of the model and the tensors it points into."""

    def _check_inputs(self, positions, cell, pairs, strain=None) -> None:
        _check_tensor("positions", positions, (self.num_atoms, 3), self.dtype, self.device)
        _check_tensor("cell", cell, (3,), self.dtype, self.device)
        if strain is not None:
            _check_tensor("strain", strain, (3, 3), self.dtype, self.device)
        _check_pairs(pairs, self.num_atoms, self.cutoff, self.device)

    def forward(self, positions: Tensor, cell: Tensor, pairs: NeighborPairs,
                strain: Tensor | None = None) -> Tensor:
        """The energy in kJ/mol of positions [N, 3] (anywhere) in the cell [3]. No host sync.

        ``strain`` is accepted only as None or zero: the virial comes from
        ``quantity.energy_force_virial``, which computes -dE/dstrain at zero strain directly.
        """
        self._check_inputs(positions, cell, pairs, strain)
        wrapped, _ = space.wrap(positions, cell)
        return self._energy_wrapped(wrapped, cell, pairs, strain)

    def _energy_wrapped(self, positions: Tensor, cell: Tensor, pairs: NeighborPairs,
                        strain: Tensor | None = None) -> Tensor:
        """``forward``'s energy for positions already in [0, L), without checks."""
        from .. import quantity

        if strain is not None and bool(strain.any()):
            raise NotImplementedError("a nonzero strain is not supported; use quantity.virial")
        return quantity._evaluate(self, positions, cell, pairs, energy=True)[0]

    def _trained_masks(self) -> dict[str, Tensor]:
        return {}


__all__ = ("Potential",)
