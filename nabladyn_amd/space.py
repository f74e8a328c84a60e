"""Orthorhombic periodic geometry on [..., 3] tensors; ``cell`` [3] holds the edge lengths in nm.

``wrap`` runs the library's HIP kernel (the reference's arithmetic, ``space.py:22-38``:
k = floor(x / L), w = x - k L, then the two ordered fixes). ``displacement`` and ``strained``
are the reference's elementwise helpers (``space.py:11-19,51-64``) for callers that need
vectors outside the fused kernels.
"""

import torch
from torch import Tensor

from . import _native


def frames(info, in_dims, args):
    """For a ``vmap`` rule that runs one frame at a time: the arguments sliced along their
    vmapped dims, one tuple per frame."""
    return [tuple(a.select(d, k) if d is not None else a for a, d in zip(args, in_dims))
            for k in range(info.batch_size)]


class _Wrap(torch.autograd.Function):
    """The wrap kernel with the reference's derivatives (``space.py:22-38``): the identity in
    the positions (the crossings are piecewise constant) and -crossings in the cell."""

    @staticmethod
    def forward(positions, cell):
        x = positions.detach().contiguous()
        out = torch.empty_like(x)
        crossings = torch.zeros(x.shape, dtype=torch.int64, device=x.device)
        c = cell.detach().to(x.dtype).contiguous()
        _native.call(
            "wrap", _native.dtype_tag(x.dtype), _native.ptr(x), _native.ptr(c),
            _native.ctypes.c_int64(x.shape[0]), _native.ptr(out), _native.ptr(crossings),
            _native.stream(x.device),
        )
        return out, crossings

    @staticmethod
    def setup_context(ctx, inputs, output):
        _, crossings = output
        ctx.mark_non_differentiable(crossings)
        ctx.save_for_backward(crossings)

    @staticmethod
    def backward(ctx, grad_wrapped, grad_crossings):
        (crossings,) = ctx.saved_tensors
        grad_cell = None
        if ctx.needs_input_grad[1]:
            grad_cell = -(grad_wrapped * crossings.to(grad_wrapped.dtype)).sum(dim=0)
        return grad_wrapped, grad_cell

    @staticmethod
    def vmap(info, in_dims, positions, cell):
        outs = [_Wrap.apply(*frame) for frame in frames(info, in_dims, (positions, cell))]
        return (torch.stack([o[0] for o in outs]), torch.stack([o[1] for o in outs])), (0, 0)


def wrap(positions: Tensor, cell: Tensor) -> tuple[Tensor, Tensor]:
    """Wrap positions [N, 3] into the cell, each coordinate in [0, L).

    Returns ``(wrapped, crossings)``, crossings [N, 3] int64 with
    ``positions == wrapped + crossings * cell`` up to rounding. GPU tensors only. Differentiable
    as the reference's: the identity in the positions, -crossings in the cell.
    """
    _native.require_gpu(positions, "positions")
    if positions.dim() != 2 or positions.shape[-1] != 3 or cell.shape != (3,):
        raise ValueError("wrap takes positions [N, 3] and a cell [3]")
    return _Wrap.apply(positions, cell)


def wrap_into(positions: Tensor, cell: Tensor, out: Tensor, images: Tensor | None = None) -> None:
    """Wrap positions [N, 3] into ``out`` (may be ``positions``), adding crossings to ``images``.

    No allocation and no sync, so a HIP graph can capture it.
    """
    _native.call(
        "wrap", _native.dtype_tag(positions.dtype), _native.ptr(positions), _native.ptr(cell),
        _native.ctypes.c_int64(positions.shape[0]), _native.ptr(out), _native.ptr(images),
        _native.stream(positions.device),
    )


def shift(positions: Tensor, delta: Tensor, cell: Tensor) -> tuple[Tensor, Tensor]:
    """Return ``positions + delta`` wrapped into the cell, with ``wrap``'s crossings."""
    return wrap(positions + delta, cell)


def _displacement_wrapped(left: Tensor, right: Tensor, cell: Tensor) -> Tensor:
    """The minimum-image vector ``left - right`` for coordinates in [0, L) (``space.py:11-19``)."""
    raw = left - right
    half_cell = 0.5 * cell
    left_shift = torch.where(raw > half_cell, cell, 0.0)
    right_shift = torch.where(raw < -half_cell, cell, 0.0)
    return (left - left_shift) - (right - right_shift)


def displacement(left: Tensor, right: Tensor, cell: Tensor) -> Tensor:
    """The minimum-image vector ``left - right`` [N, 3] in nm, for positions anywhere."""
    return _displacement_wrapped(wrap(left, cell)[0], wrap(right, cell)[0], cell)


def strained(vectors: Tensor, strain: Tensor | None) -> Tensor:
    """Deform ``vectors`` [..., 3] by (I + strain); ``None`` returns ``vectors`` itself."""
    if strain is None:
        return vectors
    return (
        vectors
        + vectors[..., 0:1] * strain[:, 0]
        + vectors[..., 1:2] * strain[:, 1]
        + vectors[..., 2:3] * strain[:, 2]
    )


__all__ = ("wrap", "shift", "displacement", "strained")
