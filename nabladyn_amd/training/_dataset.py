"""Labeled frames of one topology (``training/_dataset.py:1-105``).

The frames live on the host; ``fit`` moves what it uses to the model's device once.
"""

import numpy as np
import torch
from torch import Tensor

from ._loss import CHANNELS

# the trailing shape of each field, given F frames of N atoms
_SHAPES = {
    "positions": lambda f, n: (f, n, 3),
    "cells": lambda f, n: (f, 3),
    "energy": lambda f, n: (f,),
    "forces": lambda f, n: (f, n, 3),
    "virial": lambda f, n: (f, 3, 3),
}


def _on_host(value) -> Tensor:
    """A detached CPU copy of an array or a tensor (never a view of the caller's memory)."""
    if isinstance(value, Tensor):
        return value.detach().to("cpu", copy=True)
    return torch.from_numpy(np.array(value, copy=True))


def _validate(fields: dict[str, Tensor]) -> None:
    """The checks of ``training/_dataset.py:20-49``: shapes from the positions, one float dtype,
    finite values and positive cells; ``ValueError`` otherwise."""
    positions = fields["positions"]
    if positions.ndim != 3 or positions.shape[0] == 0 or positions.shape[2] != 3:
        raise ValueError("positions must be [F, N, 3] with at least one frame")
    frames, atoms = positions.shape[:2]
    dtype = positions.dtype
    for name, value in fields.items():
        expected = _SHAPES[name](frames, atoms)
        if tuple(value.shape) != expected:
            raise ValueError(f"{name} must be {list(expected)}, not {list(value.shape)}")
        if value.dtype != dtype or dtype not in (torch.float32, torch.float64):
            raise ValueError("every field must share one dtype, float32 or float64")
        if not bool(torch.isfinite(value).all()):
            raise ValueError(f"{name} must be finite")
    if not bool((fields["cells"] > 0).all()):
        raise ValueError("cells must be positive")


class Dataset:
    """Labeled frames of one topology, copied to the CPU.

    ``positions`` [F, N, 3] and ``cells`` [F, 3] in nm; optional labels ``energy`` [F] in
    kJ/mol, ``forces`` [F, N, 3] in kJ/mol/nm and ``virial`` [F, 3, 3] in kJ/mol, as arrays or
    tensors of one dtype, float32 or float64. Indexing with an integer, a slice or an integer
    array gives those frames as a new Dataset.

    Raises ``ValueError`` for a wrong shape or dtype, a non-finite value or a nonpositive cell.
    """

    def __init__(self, positions, cells, *, energy=None, forces=None, virial=None):
        given = (("positions", positions), ("cells", cells), ("energy", energy),
                 ("forces", forces), ("virial", virial))
        fields = {name: _on_host(value) for name, value in given if value is not None}
        _validate(fields)
        self._fields = fields

    @classmethod
    def _of(cls, fields: dict[str, Tensor]) -> "Dataset":
        selected = cls.__new__(cls)
        selected._fields = fields
        return selected

    def __len__(self) -> int:
        return int(self._fields["positions"].shape[0])

    def __getitem__(self, key) -> "Dataset":
        if isinstance(key, slice):
            rows = torch.arange(len(self))[key]
        else:
            rows = torch.as_tensor(key).cpu().reshape(-1)
        return Dataset._of({name: value[rows] for name, value in self._fields.items()})

    @property
    def positions(self) -> Tensor:
        """Positions [F, N, 3] in nm."""
        return self._fields["positions"]

    @property
    def cells(self) -> Tensor:
        """Cell lengths [F, 3] in nm."""
        return self._fields["cells"]

    @property
    def labels(self) -> dict[str, Tensor]:
        """The labels present, in channel order: ``energy``, ``forces``, ``virial``."""
        return {name: self._fields[name] for name in CHANNELS if name in self._fields}


__all__ = ("Dataset",)
