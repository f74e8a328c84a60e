"""Energy, forces and virial from one evaluation (``quantity.py``).

Each function binds a model (a LennardJones or Tabulated potential of this package, a Sum of
them, or its ``_energy_wrapped``) and returns a function of ``(positions, cell, pairs)``. Where the
reference differentiates the energy with ``torch.func``, here one fused HIP launch computes
E, F = -dE/dx and W = -dE/dstrain at zero strain (= sum over pairs of r_ij (x) F_ij) over the
pairs' columns. Results carry no autograd graph; parameter gradients come from
``nabladyn_amd.fit``.
"""

from collections.abc import Callable

import torch
from torch import Tensor

from . import _native, space
from .neighbors import NeighborPairs


def _model_of(energy_fn):
    """The Potential behind ``energy_fn`` and whether its positions still need wrapping."""
    from .potential import Potential

    if isinstance(energy_fn, Potential):
        return energy_fn, True
    owner = getattr(energy_fn, "__self__", None)
    if isinstance(owner, Potential) and getattr(energy_fn, "__name__", "") == "_energy_wrapped":
        return owner, False
    raise TypeError(
        "quantity functions take a LennardJones, Tabulated or Sum model of nabladyn_amd, "
        "or its _energy_wrapped",
    )


def _evaluate(model, positions: Tensor, cell: Tensor, pairs: NeighborPairs, *,
              energy=False, forces=False, virial=False):
    """One fused launch over ``pairs.index``; returns (E, F, W) with None for what was not asked.
    A sum takes one launch per part, each on the shortest list that reaches its cutoff
    (``potential/_sum/_potential.py:86-99``), and adds the parts' E, F and W in part order."""
    from .potential import Sum
    from .potential._sum import own_pairs

    if isinstance(model, Sum):
        parts = [_evaluate(part, positions, cell, own_pairs(pairs, part.cutoff), energy=energy,
                           forces=forces, virial=virial) for part in model.parts]
        e = torch.stack([p[0] for p in parts]).sum() if energy else None
        f = w = None
        for p in parts:
            if forces:
                f = p[1] if f is None else f + p[1]
            if virial:
                w = p[2] if w is None else w + p[2]
        return e, f, w
    dtype, device = positions.dtype, positions.device
    index = pairs.index
    if index.dtype != torch.int32 or not index.is_contiguous():
        index = index.to(torch.int32).contiguous()
    x = positions.detach().contiguous()
    cell = cell.detach().contiguous()
    pot, keep = model._native()
    capacity = index.shape[1]
    size = _native.ctypes.c_int64()
    tag = _native.dtype_tag(dtype)
    _native.call("pair_workspace_bytes", tag, _native.ctypes.byref(pot),
                 _native.ctypes.c_int64(capacity), _native.ctypes.byref(size))
    workspace = torch.empty(size.value, dtype=torch.uint8, device=device)
    e = torch.empty((), dtype=dtype, device=device) if energy else None
    f = torch.empty((x.shape[0], 3), dtype=dtype, device=device) if forces else None
    w = torch.empty((3, 3), dtype=dtype, device=device) if virial else None
    _native.call(
        "pair_efw", tag, _native.ptr(x), _native.ptr(cell), _native.ctypes.c_int64(x.shape[0]),
        _native.ptr(index), _native.ctypes.c_int64(capacity), _native.ctypes.byref(pot),
        _native.ptr(e) if e is not None else None, _native.ptr(f), _native.ptr(w),
        _native.ptr(workspace), _native.ctypes.c_int64(size.value), _native.stream(device),
    )
    del keep
    return e, f, w


def _bind(energy_fn, energy: bool, forces: bool, virial: bool):
    model, wraps = _model_of(energy_fn)

    def evaluate(positions: Tensor, cell: Tensor, pairs: NeighborPairs):
        if wraps:
            model._check_inputs(positions, cell, pairs)
            positions, _ = space.wrap(positions, cell)
        e, f, w = _evaluate(model, positions, cell, pairs, energy=energy, forces=forces, virial=virial)
        return tuple(v for v in (e, f, w) if v is not None)

    return evaluate


def force(energy_fn: Callable[..., Tensor]):
    """Bind F, the forces [N, 3] in kJ/mol/nm."""
    bound = _bind(energy_fn, False, True, False)
    return lambda positions, cell, pairs: bound(positions, cell, pairs)[0]


def energy_and_force(energy_fn: Callable[..., Tensor]):
    """Bind (E, F)."""
    return _bind(energy_fn, True, True, False)


def force_virial(energy_fn: Callable[..., Tensor]):
    """Bind (F, W)."""
    return _bind(energy_fn, False, True, True)


def virial(energy_fn: Callable[..., Tensor]):
    """Bind W, the virial [3, 3] in kJ/mol."""
    bound = _bind(energy_fn, False, False, True)
    return lambda positions, cell, pairs: bound(positions, cell, pairs)[0]


def energy_virial(energy_fn: Callable[..., Tensor]):
    """Bind (E, W)."""
    return _bind(energy_fn, True, False, True)


def energy_force_virial(energy_fn: Callable[..., Tensor]):
    """Bind (E, F, W): the energy and virial [3, 3] in kJ/mol, forces [N, 3] in kJ/mol/nm."""
    return _bind(energy_fn, True, True, True)


__all__ = ("force", "energy_and_force", "force_virial", "virial", "energy_virial",
           "energy_force_virial")
