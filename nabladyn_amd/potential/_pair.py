"""Lennard-Jones and tabulated pair potentials (``potential/_lennard_jones/_potential.py``,
``potential/_tabulated/_potential.py``): the reference's constructors, checks and parameters,
evaluated by the library's HIP kernels.
"""

import math
from collections.abc import Mapping, Sequence
from numbers import Integral, Real

import torch
from torch import Tensor, nn

from .. import _native
from ._base import Potential


def pair_index(first: Tensor, second: Tensor, num_types: int) -> Tensor:
    """The parameter row of each unordered type pair: a(2T - a - 1)/2 + b for a <= b
    (``_lennard_jones/_potential.py:143-151``)."""
    low, high = torch.minimum(first, second), torch.maximum(first, second)
    return low * (2 * num_types - low - 1) // 2 + high


def _check_cutoff(cutoff) -> None:
    if isinstance(cutoff, bool) or not isinstance(cutoff, Real) or not math.isfinite(cutoff) or cutoff <= 0:
        raise ValueError("cutoff must be a finite positive number")


def _num_types(rows: int, what: str) -> int:
    num_types = math.isqrt(2 * rows)
    if num_types == 0 or num_types * (num_types + 1) // 2 != rows:
        raise ValueError(f"{what} need T(T+1)/2 values, one per type pair")
    return num_types


def _check_atom_types(atom_types: Tensor, num_types: int) -> None:
    if (
        not isinstance(atom_types, Tensor) or atom_types.ndim != 1
        or atom_types.dtype != torch.int64 or atom_types.numel() == 0
    ):
        raise ValueError("atom_types must be a nonempty int64 vector")
    if not bool(((atom_types >= 0) & (atom_types < num_types)).all()):
        raise ValueError(f"atom_types must lie in [0, {num_types})")


def _row_mask(trainable, name: str, count: int, device) -> Tensor | None:
    """``_lennard_jones/_potential.py:92-110``."""
    if not isinstance(trainable, Mapping) or name not in trainable:
        return None
    rows = trainable[name]
    if (
        isinstance(rows, (str, bytes)) or not isinstance(rows, Sequence) or not rows
        or any(isinstance(r, bool) or not isinstance(r, Integral) or not 0 <= r < count for r in rows)
    ):
        raise ValueError(f"trainable[{name!r}] must list rows in [0, {count})")
    if len(set(rows)) == count:
        return None
    mask = torch.zeros(count, dtype=torch.bool, device=device)
    mask[[int(r) for r in rows]] = True
    return mask


class LennardJones(Potential):
    """12-6 Lennard-Jones shifted to zero at the cutoff, one sigma and epsilon per type pair."""

    def __init__(self, *, atom_types: Tensor, sigma: Tensor, epsilon: Tensor, cutoff: Real,
                 trainable=()):
        super().__init__()
        if isinstance(trainable, str) or not set(trainable) <= {"sigma", "epsilon"}:
            raise ValueError('trainable must name parameters among "sigma" and "epsilon"')
        _check_cutoff(cutoff)
        for name, value in (("sigma", sigma), ("epsilon", epsilon)):
            if not isinstance(value, Tensor) or value.ndim != 1 or value.dtype not in (torch.float32, torch.float64):
                raise ValueError(f"{name} must be a float32 or float64 vector")
        if (epsilon.shape, epsilon.dtype, epsilon.device) != (sigma.shape, sigma.dtype, sigma.device):
            raise ValueError("sigma and epsilon must share shape, dtype and device")
        num_types = _num_types(sigma.numel(), "sigma and epsilon")
        values = torch.cat([sigma, epsilon])
        if not bool((values.isfinite() & (values > 0)).all()):
            raise ValueError("sigma and epsilon must be finite and positive")
        _check_atom_types(atom_types, num_types)
        self._cutoff = float(cutoff)
        self._num_types = num_types
        self.sigma = nn.Parameter(sigma.detach().clone(), requires_grad="sigma" in trainable)
        self.epsilon = nn.Parameter(epsilon.detach().clone(), requires_grad="epsilon" in trainable)
        self.register_buffer("atom_types", atom_types.detach().to(sigma.device, copy=True))
        for name in ("sigma", "epsilon"):
            self.register_buffer(f"_{name}_rows", _row_mask(trainable, name, sigma.numel(), sigma.device),
                                 persistent=False)

    @property
    def cutoff(self) -> float:
        return self._cutoff

    @property
    def num_atoms(self) -> int:
        return self.atom_types.numel()

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
        sigma = self.sigma.detach().contiguous()
        epsilon = self.epsilon.detach().contiguous()
        types = self.atom_types.contiguous()
        pot = _native.Potential(
            kind=_native.LJ, num_types=self._num_types, atom_types=_native.ptr(types),
            sigma=_native.ptr(sigma), epsilon=_native.ptr(epsilon), table=None, start=None,
            spacing=None, knots=0, cutoff=self._cutoff,
        )
        return pot, (sigma, epsilon, types)


class Tabulated(Potential):
    """A pair potential tabulated per unordered type pair, 0 at the cutoff (Catmull-Rom)."""

    def __init__(self, *, atom_types: Tensor, table: Tensor, start: Tensor, cutoff: Real,
                 trainable=()):
        super().__init__()
        if isinstance(trainable, str) or not set(trainable) <= {"table"}:
            raise ValueError('trainable may name only "table"')
        _check_cutoff(cutoff)
        if (
            not isinstance(table, Tensor) or table.ndim != 2 or table.shape[1] < 4
            or table.dtype not in (torch.float32, torch.float64)
        ):
            raise ValueError("table must be a float32 or float64 matrix of at least 4 columns")
        if not bool(table.isfinite().all()):
            raise ValueError("table must be finite")
        if bool((table[:, -1] != 0).any()):
            raise ValueError("table must hold 0 in its last column, at the cutoff")
        num_types = _num_types(table.shape[0], "table and start")
        if not isinstance(start, Tensor) or (start.shape, start.dtype, start.device) != (
            table.shape[:1], table.dtype, table.device
        ):
            raise ValueError("start must hold one value per table row, in the table's dtype and device")
        if not bool(((start >= 0) & (start < cutoff)).all()):
            raise ValueError(f"start must lie in [0, {cutoff})")
        _check_atom_types(atom_types, num_types)
        self._cutoff = float(cutoff)
        self._num_types = num_types
        self.table = nn.Parameter(table.detach().clone(), requires_grad="table" in trainable)
        self.register_buffer("start", start.detach().clone())
        self.register_buffer("spacing", (cutoff - start.detach()) / (table.shape[1] - 1))
        self.register_buffer("atom_types", atom_types.detach().to(table.device, copy=True))
        mask = _row_mask(trainable, "table", table.shape[0], table.device)
        self.register_buffer("_table_rows", None if mask is None else mask[:, None], persistent=False)

    @property
    def cutoff(self) -> float:
        return self._cutoff

    @property
    def num_atoms(self) -> int:
        return self.atom_types.numel()

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
        table = self.table.detach().contiguous()
        start = self.start.contiguous()
        spacing = self.spacing.contiguous()
        types = self.atom_types.contiguous()
        pot = _native.Potential(
            kind=_native.TABLE, num_types=self._num_types, atom_types=_native.ptr(types),
            sigma=None, epsilon=None, table=_native.ptr(table), start=_native.ptr(start),
            spacing=_native.ptr(spacing), knots=table.shape[1], cutoff=self._cutoff,
        )
        return pot, (table, start, spacing, types)


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
