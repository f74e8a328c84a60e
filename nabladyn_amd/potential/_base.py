"""The Potential contract (``potential/_base.py``): a model of a fixed set of atoms whose one
scalar energy defines forces and virial.

A model here is an ``nn.Module`` with four properties (cutoff, atom count, dtype, device), an
energy of wrapped positions (``_energy_wrapped``), and optionally a kernel description
(``_native``) that lets ``quantity`` and the MD step evaluate it with the library's fused
kernels. ``forward`` checks its inputs, wraps the positions and returns the energy, which is
differentiable (positions, strain, parameters), as the reference's is.
"""

from abc import ABC, abstractmethod

import torch
from torch import Tensor, nn

from .. import space
from ..neighbors import NeighborPairs


def _mismatch(name, value, shape, dtype, device) -> str | None:
    """The reference's message for an input of the wrong kind, or None
    (``potential/_base.py:11-23``)."""
    if not isinstance(value, Tensor) or value.shape != shape:
        return f"{name} must be a tensor of shape {list(shape)}"
    if (value.dtype, value.device) != (dtype, device):
        return f"{name} must be {dtype} on {device} to match the model"
    return None


def _pairs_mismatch(pairs: NeighborPairs, num_atoms: int, cutoff: float, device) -> str | None:
    """The reference's message for pairs that do not fit the model, or None
    (``potential/_base.py:26-43``): every list [2, C] on the model's device and for its atoms,
    the outer list reaching the cutoff."""
    for listed in (pairs, *pairs.inner):
        shape_ok = listed.index.ndim == 2 and listed.index.shape[0] == 2
        if not shape_ok or listed.index.device != device:
            return f"pairs.index must be a [2, C] tensor on {device}"
        if listed.num_atoms != num_atoms:
            return f"the pairs are for {listed.num_atoms} atoms, the model for {num_atoms}"
    if pairs.cutoff < cutoff:
        return f"the neighbor list's cutoff {pairs.cutoff} nm is below the model's {cutoff} nm"
    return None


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
        """The cutoffs of the lists the model reads, ascending; the last is its own."""
        return (self.cutoff,)

    @property
    def _fused(self) -> bool:
        """Whether the library's fused kernels evaluate the model (``_native`` describes it)."""
        return False

    def _native(self):
        """The C-ABI description of the model and the tensors it points into."""
        raise TypeError(f"{type(self).__name__} has no kernel of its own")

    def _check_inputs(self, positions, cell, pairs, strain=None) -> None:
        problems = [
            _mismatch("positions", positions, (self.num_atoms, 3), self.dtype, self.device),
            _mismatch("cell", cell, (3,), self.dtype, self.device),
            None if strain is None else _mismatch("strain", strain, (3, 3), self.dtype, self.device),
            _pairs_mismatch(pairs, self.num_atoms, self.cutoff, self.device),
        ]
        for problem in problems:
            if problem is not None:
                raise ValueError(problem)

    def forward(self, positions: Tensor, cell: Tensor, pairs: NeighborPairs,
                strain: Tensor | None = None) -> Tensor:
        """The energy in kJ/mol of positions [N, 3] in nm (anywhere) in the cell [3].

        ``pairs`` must come from a neighbor list updated to these positions and cell, with a
        cutoff at least the model's. The optional ``strain`` [3, 3] deforms positions and cell
        together, x -> (I + strain) x. The energy is differentiable with respect to the
        positions, the strain and the model's parameters. No host sync.

        Raises ``ValueError`` for inputs of another shape, dtype or device, or for pairs of
        another system or a shorter cutoff.
        """
        self._check_inputs(positions, cell, pairs, strain)
        wrapped, _ = space.wrap(positions, cell)
        return self._energy_wrapped(wrapped, cell, pairs, strain)

    @abstractmethod
    def _energy_wrapped(self, positions: Tensor, cell: Tensor, pairs: NeighborPairs,
                        strain: Tensor | None = None) -> Tensor:
        """``forward``'s energy for positions already in [0, L), without checks."""

    def _trained_masks(self) -> dict[str, Tensor]:
        """A bool mask per parameter that trains only some of its entries (True: trained)."""
        return {}


__all__ = ("Potential",)
