"""Energy, forces and virial from one scalar energy (``quantity.py``).

Each function binds an energy — any callable ``energy(positions, cell, pairs, strain=None)``,
such as a model or its ``_energy_wrapped`` — and returns a function of
``(positions, cell, pairs)`` computing all of its results from one evaluation:

- E is a scalar in kJ/mol;
- F = -dE/dpositions is [N, 3] in kJ/mol/nm;
- W = -dE/dstrain at zero strain is [3, 3] in kJ/mol (no kinetic term); the pressure is
  (2K + trace W) / (3V).

A model the library's kernels evaluate (LJ, a table, a sum of them) or its ``_energy_wrapped``
takes one fused launch per part, which computes E, F and W together — over the list's full rows
when it keeps them (no atomics), else over its half list; its results carry no autograd graph.
Any other energy is differentiated by ``torch.func`` as the reference does
(``quantity.py:21-117``); models inside it are differentiable energies, so a user function that
adds, say, a restraint to an LJ model works unchanged.
"""

from collections.abc import Callable

import torch
from torch import Tensor

from . import space
from .neighbors import NeighborPairs


def _fused_model(energy_fn):
    """(model, whether it wraps) when the fused kernels evaluate ``energy_fn``, else None."""
    from .potential import Potential

    if isinstance(energy_fn, Potential):
        return (energy_fn, True) if energy_fn._fused else None
    owner = getattr(energy_fn, "__self__", None)
    if (isinstance(owner, Potential) and owner._fused
            and getattr(energy_fn, "__name__", "") == "_energy_wrapped"):
        return owner, False
    return None


def _evaluate(model, positions: Tensor, cell: Tensor, pairs: NeighborPairs, *,
              energy=False, forces=False, virial=False):
    """E, F, W of a fused model on wrapped positions, None for what was not asked: one launch
    per part, each part on the shortest list that reaches its cutoff, added in part order
    (``potential/_sum/_potential.py:86-99``)."""
    from .potential import Sum
    from .potential._energy import evaluate, evaluate_full
    from .potential._sum import list_for

    if isinstance(model, Sum):
        results = [_evaluate(part, positions, cell, list_for(pairs, part.cutoff), energy=energy,
                             forces=forces, virial=virial) for part in model.parts]
        summed = []
        for k in range(3):
            values = [r[k] for r in results]
            if values[0] is None:
                summed.append(None)
            elif k == 0:
                summed.append(torch.stack(values).sum())
            else:
                summed.append(_add(values))
        return tuple(summed)
    if pairs.full is not None:
        return evaluate_full(model, positions, cell, pairs.full, energy, forces, virial)
    index = pairs.index
    if index.dtype != torch.int32 or not index.is_contiguous():
        index = index.to(torch.int32).contiguous()
    return evaluate(model, positions, cell, index, None, energy, forces, virial)


def _add(values):
    total = values[0]
    for value in values[1:]:
        total = total + value
    return total


def _with_strain(energy_fn):
    """``energy_fn`` with the strain as its second argument, for ``torch.func``
    (``quantity.py:21-32``)."""

    def strained_energy(positions, strain, cell, pairs):
        return energy_fn(positions, cell, pairs, strain=strain)

    return strained_energy


def _bind(energy_fn, energy: bool, forces: bool, virial: bool):
    fused = _fused_model(energy_fn)
    if fused is not None:
        model, wraps = fused

        def evaluate(positions: Tensor, cell: Tensor, pairs: NeighborPairs):
            if wraps:
                model._check_inputs(positions, cell, pairs)
                positions, _ = space.wrap(positions.detach(), cell)
            results = _evaluate(model, positions, cell, pairs, energy=energy, forces=forces,
                                virial=virial)
            return tuple(v for v in results if v is not None)

        return evaluate

    # any other energy: the reference's route, one grad_and_value over positions and strain
    argnums = tuple(k for k, wanted in ((0, forces), (1, virial)) if wanted)
    transform = torch.func.grad_and_value(_with_strain(energy_fn),
                                          argnums=argnums if len(argnums) > 1 else argnums[0])

    def evaluate(positions: Tensor, cell: Tensor, pairs: NeighborPairs):
        gradients, value = transform(positions, positions.new_zeros((3, 3)), cell, pairs)
        if len(argnums) == 1:
            gradients = (gradients,)
        out = [value] if energy else []
        out.extend(-g for g in gradients)
        return tuple(out)

    return evaluate


def force(energy_fn: Callable[..., Tensor]):
    """Bind F, the forces [N, 3] in kJ/mol/nm."""
    bound = _bind(energy_fn, False, True, False)
    return lambda positions, cell, pairs: bound(positions, cell, pairs)[0]


def energy_and_force(energy_fn: Callable[..., Tensor]):
    """Bind (E, F), the energy in kJ/mol and the forces [N, 3] in kJ/mol/nm."""
    return _bind(energy_fn, True, True, False)


def force_virial(energy_fn: Callable[..., Tensor]):
    """Bind (F, W), the forces [N, 3] in kJ/mol/nm and the virial [3, 3] in kJ/mol."""
    return _bind(energy_fn, False, True, True)


def virial(energy_fn: Callable[..., Tensor]):
    """Bind W, the virial [3, 3] in kJ/mol."""
    bound = _bind(energy_fn, False, False, True)
    return lambda positions, cell, pairs: bound(positions, cell, pairs)[0]


def energy_virial(energy_fn: Callable[..., Tensor]):
    """Bind (E, W), the energy in kJ/mol and the virial [3, 3] in kJ/mol."""
    return _bind(energy_fn, True, False, True)


def energy_force_virial(energy_fn: Callable[..., Tensor]):
    """Bind (E, F, W): the energy and virial [3, 3] in kJ/mol, forces [N, 3] in kJ/mol/nm.

    ``energy_fn`` is a model, or any energy with its call signature. The result is a function of
    ``(positions, cell, pairs)``, with ``pairs`` from a ``NeighborList``, that returns all three
    from one evaluation. F = -dE/dr in the positions r; W = -dE/de in the strain e, at zero
    strain, with no kinetic term, so the pressure is (2K + tr W) / (3V).
    """
    return _bind(energy_fn, True, True, True)


__all__ = ("force", "energy_and_force", "force_virial", "virial", "energy_virial",
           "energy_force_virial")
