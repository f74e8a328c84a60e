"""Lennard-Jones and tabulated pair potentials (``potential/_lennard_jones/_potential.py``,
``potential/_tabulated/_potential.py``, built by ``potential/_api.py``).

Both are ``_PairModel`` s: one parameter row per unordered type pair (``pair_index``), atom types
as a buffer, and an energy that is one launch of the library's pair kernel (``_energy``), with
the derivatives that launch and the parameter-gradient kernel provide. Each subclass states its
parameters, how the C ABI describes it, and the shape of its parameter gradient.
"""

from numbers import Real

import torch
from torch import Tensor, nn

from .. import _native
from ..neighbors import NeighborPairs
from . import _parameters as check
from ._base import Potential
from ._energy import pair_energy


def pair_index(first: Tensor, second: Tensor, num_types: int) -> Tensor:
    """The parameter row of each unordered type pair: a(2T - a - 1)/2 + b for a <= b
    (``_lennard_jones/_potential.py:143-151``)."""
    low, high = torch.minimum(first, second), torch.maximum(first, second)
    return low * (2 * num_types - low - 1) // 2 + high


class _PairModel(Potential):
    """What LJ and tables share: cutoff, types, the fused energy and its derivatives."""

    def __init__(self, cutoff: float, num_types: int, atom_types: Tensor, device):
        super().__init__()
        self._cutoff = cutoff
        self._num_types = num_types
        self.register_buffer("atom_types", atom_types.detach().to(device, copy=True))

    @property
    def cutoff(self) -> float:
        return self._cutoff

    @property
    def num_atoms(self) -> int:
        return self.atom_types.numel()

    @property
    def _fused(self) -> bool:
        return True

    def _energy_wrapped(self, positions: Tensor, cell: Tensor, pairs: NeighborPairs,
                        strain: Tensor | None = None) -> Tensor:
        return pair_energy(self, positions, cell, pairs.index, strain)

    def _describe(self, **fields) -> tuple:
        """The C-ABI ``ndn_potential`` of the model and the tensors it points into."""
        types = self.atom_types.contiguous()
        tensors = {name: value.detach().contiguous() for name, value in fields.items()
                   if isinstance(value, Tensor)}
        pot = _native.Potential(
            kind=fields["kind"], num_types=self._num_types, atom_types=_native.ptr(types),
            sigma=_native.ptr(tensors.get("sigma")), epsilon=_native.ptr(tensors.get("epsilon")),
            table=_native.ptr(tensors.get("table")), start=_native.ptr(tensors.get("start")),
            spacing=_native.ptr(tensors.get("spacing")), knots=fields.get("knots", 0),
            cutoff=self._cutoff)
        return pot, (types, *tensors.values())


class LennardJones(_PairModel):
    """12-6 Lennard-Jones shifted to zero at the cutoff, one sigma and epsilon per type pair."""

    def __init__(self, *, atom_types: Tensor, sigma: Tensor, epsilon: Tensor, cutoff: Real,
                 trainable=()):
        names = check.trainable_names(
            trainable, ("sigma", "epsilon"),
            'trainable must name parameters among "sigma" and "epsilon"')
        cutoff = check.cutoff_value(cutoff)
        num_types = check.lj_parameters(sigma, epsilon)
        check.atom_types_for(atom_types, num_types)
        super().__init__(cutoff, num_types, atom_types, sigma.device)
        self.sigma = nn.Parameter(sigma.detach().clone(), requires_grad="sigma" in names)
        self.epsilon = nn.Parameter(epsilon.detach().clone(), requires_grad="epsilon" in names)
        for name in ("sigma", "epsilon"):
            mask = check.trained_rows(trainable, name, sigma.numel(), sigma.device)
            self.register_buffer(f"_{name}_rows", mask, persistent=False)

    @property
    def dtype(self) -> torch.dtype:
        return self.sigma.dtype

    @property
    def device(self) -> torch.device:
        return self.sigma.device

    def _trained_masks(self) -> dict[str, Tensor]:
        masks = {"sigma": self._sigma_rows, "epsilon": self._epsilon_rows}
        return {name: mask for name, mask in masks.items() if mask is not None}

    def _native(self):
        return self._describe(kind=_native.LJ, sigma=self.sigma, epsilon=self.epsilon)

    def _differentiable_parameters(self):
        return self.sigma, self.epsilon

    def _parameter_gradient_shape(self) -> tuple[int, ...]:
        return (2, self.sigma.numel())

    @staticmethod
    def _split_parameter_gradient(gradient: Tensor):
        return gradient[0], gradient[1]


class Tabulated(_PairModel):
    """A pair potential tabulated per unordered type pair, 0 at the cutoff (Catmull-Rom)."""

    def __init__(self, *, atom_types: Tensor, table: Tensor, start: Tensor, cutoff: Real,
                 trainable=()):
        names = check.trainable_names(trainable, ("table",), 'trainable may name only "table"')
        cutoff = check.cutoff_value(cutoff)
        num_types = check.table_parameters(table, start, cutoff)
        check.atom_types_for(atom_types, num_types)
        super().__init__(cutoff, num_types, atom_types, table.device)
        self.table = nn.Parameter(table.detach().clone(), requires_grad="table" in names)
        self.register_buffer("start", start.detach().clone())
        # spacing as the reference computes it, in the table's dtype
        self.register_buffer("spacing", (cutoff - start.detach()) / (table.shape[1] - 1))
        mask = check.trained_rows(trainable, "table", table.shape[0], table.device)
        self.register_buffer("_table_rows", None if mask is None else mask[:, None],
                             persistent=False)

    @property
    def dtype(self) -> torch.dtype:
        return self.table.dtype

    @property
    def device(self) -> torch.device:
        return self.table.device

    @property
    def knots(self) -> int:
        return self.table.shape[1]

    def _trained_masks(self) -> dict[str, Tensor]:
        return {} if self._table_rows is None else {"table": self._table_rows}

    def _native(self):
        return self._describe(kind=_native.TABLE, table=self.table, start=self.start,
                              spacing=self.spacing, knots=self.table.shape[1])

    def _differentiable_parameters(self):
        return self.table, None

    def _parameter_gradient_shape(self) -> tuple[int, ...]:
        return tuple(self.table.shape)

    @staticmethod
    def _split_parameter_gradient(gradient: Tensor):
        return gradient, None


def lennard_jones(*, atom_types: Tensor, sigma: Tensor, epsilon: Tensor, cutoff: Real,
                  trainable=()) -> Potential:
    """Build a 12-6 LJ model shifted to zero at the cutoff (``potential/_api.py:18-57``)."""
    return LennardJones(atom_types=atom_types, sigma=sigma, epsilon=epsilon, cutoff=cutoff,
                        trainable=trainable)


def tabulated(*, atom_types: Tensor, table: Tensor, start: Tensor, cutoff: Real,
              trainable=()) -> Potential:
    """Build a tabulated pair model, 0 at the cutoff (``potential/_api.py:296-331``)."""
    return Tabulated(atom_types=atom_types, table=table, start=start, cutoff=cutoff,
                     trainable=trainable)


__all__ = ("LennardJones", "Tabulated", "lennard_jones", "tabulated", "pair_index")
