"""The energy of an LJ or tabulated model as a differentiable tensor.

``pair_energy`` returns E of a model on a half list from one launch of the library's pair
kernel, which computes E, F = -dE/dx and W = -dE/dstrain together. Its autograd function hands
F and W back in the backward pass (-F gE for the positions, -W gE for the strain) and launches
the parameter-gradient kernel only when a parameter asks for its gradient. So
``torch.func.grad(model)``, ``torch.autograd.grad(model(...), parameters)`` and the
``quantity`` functions of a plain callable that calls a model compose as they do over the
reference's eager energies (``potential/_base.py:82-106``, ``quantity.py:93-117``).

Limits, stated where they apply:
- first derivatives only: a derivative of the backward's results (forces in a loss) raises;
  force and virial fitting has its own hand-written backward (``nabladyn_amd.training``);
- no derivative with respect to the cell (the reference's is that of its minimum-image
  shifts); the cell's gradient is None;
- ``vmap`` over frames runs one launch per frame.
"""

import torch
from torch import Tensor

from .. import _native
from ..space import frames as _frames


def _workspace(model, capacity: int, device, native=None):
    pot, keep = model._native() if native is None else native
    size = _native.ctypes.c_int64()
    _native.call("pair_workspace_bytes", _native.dtype_tag(model.dtype), _native.ctypes.byref(pot),
                 _native.ctypes.c_int64(capacity), _native.ctypes.byref(size))
    return torch.empty(size.value, dtype=torch.uint8, device=device), pot, keep


def evaluate(model, positions: Tensor, cell: Tensor, index: Tensor, strain: Tensor | None,
             energy=True, forces=True, virial=True, native=None):
    """E [], F [N, 3] and W [3, 3] of wrapped positions on the half list ``index`` [2, C] (int32,
    contiguous), at ``strain`` or none, from one launch; None for what is not asked. ``native``
    is a description from ``model._native()`` taken earlier, else the model's own now."""
    x = positions.detach().contiguous()
    cell = cell.detach().contiguous()
    dtype, device = x.dtype, x.device
    capacity = index.shape[1]
    workspace, pot, keep = _workspace(model, capacity, device, native)
    e = torch.empty((), dtype=dtype, device=device) if energy else None
    f = torch.empty((x.shape[0], 3), dtype=dtype, device=device) if forces else None
    w = torch.empty((3, 3), dtype=dtype, device=device) if virial else None
    common = (_native.dtype_tag(dtype), _native.ptr(x), _native.ptr(cell),
              _native.ctypes.c_int64(x.shape[0]), _native.ptr(index),
              _native.ctypes.c_int64(capacity), _native.ctypes.byref(pot))
    outputs = (_native.ptr(e), _native.ptr(f), _native.ptr(w), _native.ptr(workspace),
               _native.ctypes.c_int64(workspace.numel()), _native.stream(device))
    if strain is None:
        _native.call("pair_efw", *common, *outputs)
    else:
        _native.call("pair_efw_strained", *common,
                     _native.ptr(strain.detach().to(dtype).contiguous()), *outputs)
    del keep
    return e, f, w


def evaluate_full(model, positions: Tensor, cell: Tensor, full, energy=True, forces=True,
                  virial=True):
    """E, F and W over a list's full rows (``NeighborPairs.full``): one launch, each atom's
    force summed in registers, no atomics; None for what is not asked."""
    x = positions.detach().contiguous()
    cell = cell.detach().contiguous()
    dtype, device = x.dtype, x.device
    workspace, pot, keep = _workspace(model, 1, device)
    e = torch.empty((), dtype=dtype, device=device) if energy else None
    f = torch.empty((x.shape[0], 3), dtype=dtype, device=device) if forces else None
    w = torch.empty((3, 3), dtype=dtype, device=device) if virial else None
    _native.call("pair_efw_full", _native.dtype_tag(dtype), _native.ptr(x), _native.ptr(cell),
                 _native.ctypes.c_int64(x.shape[0]), _native.ptr(full.rows),
                 _native.ptr(full.counts), _native.ptr(full.slices), _native.ptr(full.index),
                 _native.ctypes.c_int64(full.index.numel()), _native.ctypes.byref(pot),
                 _native.ptr(e), _native.ptr(f), _native.ptr(w), _native.ptr(workspace),
                 _native.ctypes.c_int64(workspace.numel()), _native.stream(device))
    del keep
    return e, f, w


def parameter_gradient(model, positions: Tensor, cell: Tensor, index: Tensor,
                       strain: Tensor | None) -> Tensor:
    """dE/dparameters in fp64: LJ [2, P] (sigma, epsilon), a table [P, K]."""
    x = positions.detach().contiguous()
    cell = cell.detach().contiguous()
    capacity = index.shape[1]
    workspace, pot, keep = _workspace(model, capacity, x.device)
    shape = model._parameter_gradient_shape()
    out = torch.empty(shape, dtype=torch.float64, device=x.device)
    s = None if strain is None else strain.detach().to(x.dtype).contiguous()
    _native.call("pair_parameter_gradient", _native.dtype_tag(x.dtype), _native.ptr(x),
                 _native.ptr(cell), _native.ctypes.c_int64(x.shape[0]), _native.ptr(index),
                 _native.ctypes.c_int64(capacity), _native.ctypes.byref(pot), _native.ptr(s),
                 _native.ptr(out), _native.ctypes.c_int64(out.numel()), _native.ptr(workspace),
                 _native.ctypes.c_int64(workspace.numel()), _native.stream(x.device))
    del keep
    return out


def parameter_vjp(model, positions: Tensor, cell: Tensor, index: Tensor, grad_energy,
                  grad_forces, grad_virial, native=None) -> Tensor:
    """gE dE/dparameters + gF . dF/dparameters + gW : dW/dparameters in fp64, the layout of
    ``parameter_gradient``; each upstream gradient may be None (zero). ``native`` as for
    ``evaluate``."""
    x = positions.detach().contiguous()
    cell = cell.detach().contiguous()
    capacity = index.shape[1]
    workspace, pot, keep = _workspace(model, capacity, x.device, native)
    out = torch.empty(model._parameter_gradient_shape(), dtype=torch.float64, device=x.device)

    def upstream(g):
        return None if g is None else g.detach().to(x.dtype).contiguous()

    ge, gf, gw = upstream(grad_energy), upstream(grad_forces), upstream(grad_virial)
    _native.call("pair_parameter_vjp", _native.dtype_tag(x.dtype), _native.ptr(x),
                 _native.ptr(cell), _native.ctypes.c_int64(x.shape[0]), _native.ptr(index),
                 _native.ctypes.c_int64(capacity), _native.ctypes.byref(pot), _native.ptr(ge),
                 _native.ptr(gf), _native.ptr(gw), _native.ptr(out),
                 _native.ctypes.c_int64(out.numel()), _native.ptr(workspace),
                 _native.ctypes.c_int64(workspace.numel()), _native.stream(x.device))
    del keep
    return out


def _frames_workspace(model, frames: int, capacity: int, device, native):
    pot, keep = native
    size = _native.ctypes.c_int64()
    _native.call("pair_frames_workspace_bytes", _native.dtype_tag(model.dtype),
                 _native.ctypes.byref(pot), _native.ctypes.c_int64(frames),
                 _native.ctypes.c_int64(capacity), _native.ctypes.byref(size))
    return torch.empty(size.value, dtype=torch.uint8, device=device), pot, keep


def evaluate_frames(model, positions: Tensor, cells: Tensor, stacks: Tensor, native):
    """E [F], F [F, N, 3] and W [F, 3, 3] of F frames of wrapped positions [F, N, 3] in cells
    [F, 3] over pair stacks [F, 2, C] (int32, contiguous), from one launch; ``native`` is
    ``model._native()``."""
    x = positions.detach().contiguous()
    cells = cells.detach().contiguous()
    frames, atoms = x.shape[0], x.shape[1]
    dtype, device = x.dtype, x.device
    workspace, pot, keep = _frames_workspace(model, frames, stacks.shape[2], device, native)
    e = torch.empty(frames, dtype=dtype, device=device)
    f = torch.empty((frames, atoms, 3), dtype=dtype, device=device)
    w = torch.empty((frames, 3, 3), dtype=dtype, device=device)
    _native.call("pair_efw_frames", _native.dtype_tag(dtype), _native.ptr(x), _native.ptr(cells),
                 _native.ctypes.c_int64(frames), _native.ctypes.c_int64(atoms),
                 _native.ptr(stacks), _native.ctypes.c_int64(stacks.shape[2]),
                 _native.ctypes.byref(pot), _native.ptr(e), _native.ptr(f), _native.ptr(w),
                 _native.ptr(workspace), _native.ctypes.c_int64(workspace.numel()),
                 _native.stream(device))
    del keep
    return e, f, w


def evaluate_frames_full(model, positions: Tensor, cells: Tensor, rows, native):
    """``evaluate_frames`` over full rows (``rows``: counts [F, N], slices [F, N / 64 + 1] and
    index [F, capacity], int32, the layout ``training._loss.frame_rows`` builds): a lane per
    row, no atomics."""
    x = positions.detach().contiguous()
    cells = cells.detach().contiguous()
    frames, atoms = x.shape[0], x.shape[1]
    dtype, device = x.dtype, x.device
    workspace, pot, keep = _frames_workspace(model, frames, atoms, device, native)
    e = torch.empty(frames, dtype=dtype, device=device)
    f = torch.empty((frames, atoms, 3), dtype=dtype, device=device)
    w = torch.empty((frames, 3, 3), dtype=dtype, device=device)
    counts, slices, index = rows
    _native.call("pair_efw_full_frames", _native.dtype_tag(dtype), _native.ptr(x),
                 _native.ptr(cells), _native.ctypes.c_int64(frames), _native.ctypes.c_int64(atoms),
                 _native.ptr(counts), _native.ptr(slices), _native.ptr(index),
                 _native.ctypes.c_int64(index.shape[1]), _native.ctypes.byref(pot), _native.ptr(e),
                 _native.ptr(f), _native.ptr(w), _native.ptr(workspace),
                 _native.ctypes.c_int64(workspace.numel()), _native.stream(device))
    del keep
    return e, f, w


def parameter_vjp_frames(model, positions: Tensor, cells: Tensor, stacks: Tensor, grad_energy,
                         grad_forces, grad_virial, native) -> Tensor:
    """``parameter_vjp`` summed over the frames of ``evaluate_frames``; upstream gradients
    [F], [F, N, 3], [F, 3, 3], each may be None."""
    x = positions.detach().contiguous()
    cells = cells.detach().contiguous()
    frames, atoms = x.shape[0], x.shape[1]
    workspace, pot, keep = _frames_workspace(model, frames, stacks.shape[2], x.device, native)
    out = torch.empty(model._parameter_gradient_shape(), dtype=torch.float64, device=x.device)

    def upstream(g):
        return None if g is None else g.detach().to(x.dtype).contiguous()

    ge, gf, gw = upstream(grad_energy), upstream(grad_forces), upstream(grad_virial)
    _native.call("pair_parameter_vjp_frames", _native.dtype_tag(x.dtype), _native.ptr(x),
                 _native.ptr(cells), _native.ctypes.c_int64(frames), _native.ctypes.c_int64(atoms),
                 _native.ptr(stacks), _native.ctypes.c_int64(stacks.shape[2]),
                 _native.ctypes.byref(pot), _native.ptr(ge), _native.ptr(gf), _native.ptr(gw),
                 _native.ptr(out), _native.ctypes.c_int64(out.numel()), _native.ptr(workspace),
                 _native.ctypes.c_int64(workspace.numel()), _native.stream(x.device))
    del keep
    return out


class _FirstOrderOnly(torch.autograd.Function):
    """Passes ``value`` through; differentiating it again raises. It marks the backward's
    results, which are exact first derivatives but carry no second-derivative graph."""

    generate_vmap_rule = True

    @staticmethod
    def forward(value, *dependencies):
        return value.clone()

    @staticmethod
    def setup_context(ctx, inputs, output):
        pass

    @staticmethod
    def backward(ctx, grad):
        raise RuntimeError(
            "second derivatives of the LJ and tabulated energies are not available (the fused "
            "kernel returns F and W, not their derivatives); force and virial fitting uses "
            "nabladyn_amd.training, whose backward is the parameter VJP kernel")


class _PairEnergy(torch.autograd.Function):
    """E from one launch; F and W kept for the backward pass (see the module docstring)."""

    @staticmethod
    def forward(positions, cell, strain, first, second, model, index):
        e, f, w = evaluate(model, positions, cell, index, strain)
        return e, f, w

    @staticmethod
    def setup_context(ctx, inputs, output):
        positions, cell, strain, first, second, model, index = inputs
        _, f, w = output
        ctx.mark_non_differentiable(f, w)
        ctx.model, ctx.index, ctx.strained = model, index, strain is not None
        ctx.save_for_backward(positions, cell, strain if strain is not None else positions.new_empty(0),
                              first, second if second is not None else positions.new_empty(0), f, w)

    @staticmethod
    def backward(ctx, grad_energy, grad_forces, grad_virial):
        positions, cell, strain, first, second, f, w = ctx.saved_tensors
        needs = ctx.needs_input_grad
        strain_arg = strain if ctx.strained else None
        live = [t for t in (positions, first, second) if t.requires_grad]
        grad_positions = grad_strain = grad_first = grad_second = None
        if needs[0]:
            grad_positions = _FirstOrderOnly.apply(-f * grad_energy, *live)
        if needs[2]:
            grad_strain = _FirstOrderOnly.apply(-w * grad_energy, *live)
        if needs[3] or needs[4]:
            g = parameter_gradient(ctx.model, positions, cell, ctx.index, strain_arg)
            g = g.to(positions.dtype) * grad_energy
            parts = ctx.model._split_parameter_gradient(g)
            if needs[3]:
                grad_first = _FirstOrderOnly.apply(parts[0], *live)
            if needs[4]:
                grad_second = _FirstOrderOnly.apply(parts[1], *live)
        return grad_positions, None, grad_strain, grad_first, grad_second, None, None

    @staticmethod
    def vmap(info, in_dims, positions, cell, strain, first, second, model, index):
        args = (positions, cell, strain, first, second)
        dims = tuple(in_dims[:5])
        outs = [_PairEnergy.apply(*frame, model, index) for frame in _frames(info, dims, args)]
        stacked = tuple(torch.stack([o[k] for o in outs]) for k in range(3))
        return stacked, (0, 0, 0)


def pair_energy(model, positions: Tensor, cell: Tensor, index: Tensor,
                strain: Tensor | None) -> Tensor:
    """E of `model` (LJ or tabulated) on wrapped positions and a half list, differentiable with
    respect to positions, strain and the model's parameters."""
    if index.dtype != torch.int32 or not index.is_contiguous():
        index = index.to(torch.int32).contiguous()
    first, second = model._differentiable_parameters()
    e, _, _ = _PairEnergy.apply(positions, cell, strain, first, second, model, index)
    return e
