"""Validation of a pair model's constructor arguments (``potential/_api.py:18-57,296-331``).

The error types and messages are the reference's (a caller may match them); each check here is
one predicate over the argument and raises ``ValueError`` with that message.
"""

import math
from collections.abc import Mapping, Sequence
from numbers import Integral, Real

import torch
from torch import Tensor

FLOATS = (torch.float32, torch.float64)


def cutoff_value(cutoff) -> float:
    ok = isinstance(cutoff, Real) and not isinstance(cutoff, bool) and math.isfinite(cutoff)
    if not ok or cutoff <= 0:
        raise ValueError("cutoff must be a finite positive number")
    return float(cutoff)


def types_from_rows(rows: int, what: str) -> int:
    """T from T(T + 1) / 2 parameter rows, one per unordered type pair."""
    t = math.isqrt(2 * rows)
    if t < 1 or t * (t + 1) // 2 != rows:
        raise ValueError(f"{what} need T(T+1)/2 values, one per type pair")
    return t


def atom_types_for(atom_types, num_types: int) -> Tensor:
    ok = (isinstance(atom_types, Tensor) and atom_types.ndim == 1
          and atom_types.dtype == torch.int64 and atom_types.numel() > 0)
    if not ok:
        raise ValueError("atom_types must be a nonempty int64 vector")
    if not bool(((atom_types >= 0) & (atom_types < num_types)).all()):
        raise ValueError(f"atom_types must lie in [0, {num_types})")
    return atom_types


def trainable_names(trainable, allowed: tuple[str, ...], message: str) -> set[str]:
    """The parameter names ``trainable`` selects (a tuple of names, or a mapping of names to
    trained rows)."""
    names = set(trainable) if not isinstance(trainable, str) else None
    if names is None or not names <= set(allowed):
        raise ValueError(message)
    return names


def trained_rows(trainable, name: str, count: int, device) -> Tensor | None:
    """The bool mask [count] of the rows of parameter `name` that train, or None when all do
    (``_lennard_jones/_potential.py:92-110``): ``trainable[name]`` lists row indices."""
    if not isinstance(trainable, Mapping) or name not in trainable:
        return None
    rows = trainable[name]
    valid = (isinstance(rows, Sequence) and not isinstance(rows, (str, bytes)) and len(rows) > 0
             and all(isinstance(r, Integral) and not isinstance(r, bool) and 0 <= r < count
                     for r in rows))
    if not valid:
        raise ValueError(f"trainable[{name!r}] must list rows in [0, {count})")
    if len(set(rows)) == count:
        return None
    mask = torch.zeros(count, dtype=torch.bool, device=device)
    mask[torch.tensor(sorted(set(int(r) for r in rows)), device=device)] = True
    return mask


def float_vector(name: str, value) -> Tensor:
    if not isinstance(value, Tensor) or value.ndim != 1 or value.dtype not in FLOATS:
        raise ValueError(f"{name} must be a float32 or float64 vector")
    return value


def lj_parameters(sigma, epsilon) -> int:
    """Checks sigma and epsilon; returns the number of atom types."""
    float_vector("sigma", sigma)
    float_vector("epsilon", epsilon)
    if (epsilon.shape, epsilon.dtype, epsilon.device) != (sigma.shape, sigma.dtype, sigma.device):
        raise ValueError("sigma and epsilon must share shape, dtype and device")
    num_types = types_from_rows(sigma.numel(), "sigma and epsilon")
    both = torch.cat([sigma.detach(), epsilon.detach()])
    if not bool((both.isfinite() & (both > 0)).all()):
        raise ValueError("sigma and epsilon must be finite and positive")
    return num_types


def table_parameters(table, start, cutoff: float) -> int:
    """Checks a table [P, K >= 4] (0 in its last column) and its start [P]; returns the number
    of atom types (``_tabulated/_potential.py:22-56``)."""
    ok = (isinstance(table, Tensor) and table.ndim == 2 and table.shape[1] >= 4
          and table.dtype in FLOATS)
    if not ok:
        raise ValueError("table must be a float32 or float64 matrix of at least 4 columns")
    if not bool(table.isfinite().all()):
        raise ValueError("table must be finite")
    if bool((table[:, -1] != 0).any()):
        raise ValueError("table must hold 0 in its last column, at the cutoff")
    num_types = types_from_rows(table.shape[0], "table and start")
    if not isinstance(start, Tensor) or (start.shape, start.dtype, start.device) != (
            table.shape[:1], table.dtype, table.device):
        raise ValueError("start must hold one value per table row, in the table's dtype and device")
    if not bool(((start >= 0) & (start < cutoff)).all()):
        raise ValueError(f"start must lie in [0, {cutoff})")
    return num_types
