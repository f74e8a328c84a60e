"""Force-and-virial matching of a tabulated pair potential with a hand-written backward.

``TableFrames`` holds F frames of one system (positions [F, N, 3], cells [F, 3]) and a one-type
``Tabulated`` model; ``predict(table)`` returns energy [F], forces [F, N, 3] and virial [F, 3, 3]
through ``TabulatedEFW``, a ``torch.autograd.Function`` whose forward is the batched HIP engine
(the frames are its systems) and whose backward is ``ndn_batch_table_vjp``: E, F and W are linear
in the knots, so the vector-Jacobian product is one pass over the pairs, exact, with no second
derivative graph. ``loss`` and ``scales`` restate the reference's supervised loss
(``training/_loss.py:26-125``): each channel's mean of ((X - X*) / s)^2 per frame, averaged over
frames and weighted.
"""

import torch
from torch import Tensor

from . import _native
from .batch import Batch, padded_table
from .potential import Tabulated


def scales(energy: Tensor | None = None, forces: Tensor | None = None,
           virial: Tensor | None = None) -> dict[str, float]:
    """Channel scales from training labels (``training/_loss.py:26-44``): energy and virial the RMS
    about the mean over frames, forces the plain RMS; constant labels fall back to the plain RMS,
    all-zero labels to 1. Syncs."""
    result = {}
    for name, labels, about_mean in (("energy", energy, True), ("forces", forces, False),
                                     ("virial", virial, True)):
        if labels is not None:
            rms = float(labels.square().mean().sqrt())
            spread = rms
            if about_mean:
                spread = float((labels - labels.mean(dim=0)).square().mean().sqrt())
            result[name] = spread or rms or 1.0
    return result


def loss(predicted: dict[str, Tensor], labels: dict[str, Tensor], weights: dict[str, float],
         channel_scales: dict[str, float]) -> tuple[Tensor, dict[str, Tensor]]:
    """The weighted sum over channels of each frame's mean ((X - X*) / s)^2, averaged over frames
    (``training/_loss.py:105-125`` with no coefficients)."""
    terms = {}
    for name, value in predicted.items():
        squares = ((value - labels[name]) / channel_scales[name]).square()
        terms[name] = squares.reshape(squares.shape[0], -1).mean(dim=1).mean()
    return sum(weights[name] * term for name, term in terms.items()), terms


class TableFrames:
    """Frames of one system under a one-type tabulated model, evaluated as one batch."""

    def __init__(self, model: Tabulated, positions: Tensor, cells: Tensor, *,
                 schedule=("wave", 2)):
        if not isinstance(model, Tabulated) or model._num_types != 1:
            raise TypeError("TableFrames takes a one-type Tabulated model")
        if positions.ndim != 3 or positions.shape[1] != model.num_atoms or positions.shape[2] != 3:
            raise ValueError("positions must be [F, N, 3] with the model's N")
        if cells.shape != (positions.shape[0], 3):
            raise ValueError("cells must be [F, 3]")
        self.model = model
        self.frames, self.num_atoms = positions.shape[0], positions.shape[1]
        systems = [(positions[f], cells[f], model) for f in range(self.frames)]
        self.batch = Batch(systems, schedule=schedule)
        self.schedule = tuple(schedule)
        size = int(_vjp_bytes()(_native.ctypes.c_int64(self.batch.num_atoms),
                                _native.ctypes.c_int64(model.knots)))
        self.workspace = torch.empty(size, dtype=torch.uint8, device=positions.device)

    def set_table(self, table: Tensor) -> None:
        """Pads `table` [1, K] into the batch's table row, in place."""
        padded, _, _ = padded_table(_TableView(table, self.model), self.batch.dtype)
        self.batch.tables[0].copy_(padded)

    def predict(self, table: Tensor) -> dict[str, Tensor]:
        """Energy [F], forces [F, N, 3], virial [F, 3, 3] at `table`, differentiable in it."""
        energy, forces, virial = TabulatedEFW.apply(table, self)
        return {"energy": energy, "forces": forces, "virial": virial}


class _TableView:
    """Duck-typed model for padded_table: the table under training, the model's start/spacing."""

    def __init__(self, table, model):
        self.table, self.start, self.spacing = table, model.start, model.spacing


def _vjp_bytes():
    fn = _native.lib().ndn_batch_vjp_bytes
    fn.restype = _native.ctypes.c_int64
    return fn


class TabulatedEFW(torch.autograd.Function):
    """(table [1, K], TableFrames) -> (E [F], F [F, N, 3], W [F, 3, 3]); backward: d/dtable."""

    @staticmethod
    def forward(table: Tensor, frames: TableFrames):
        frames.set_table(table.detach())
        energy, forces, virial = frames.batch.evaluate() if frames.schedule is None else _launch(frames)
        f = frames.frames
        return (energy.clone(), forces.reshape(f, frames.num_atoms, 3).clone(), virial.clone())

    @staticmethod
    def setup_context(ctx, inputs, output):
        table, frames = inputs
        ctx.frames = frames
        ctx.table_shape = table.shape
        ctx.table_dtype = table.dtype

    @staticmethod
    def backward(ctx, grad_energy, grad_forces, grad_virial):
        frames = ctx.frames
        b = frames.batch
        dtype, device = b.dtype, b.device
        f = frames.frames

        def dense(g, shape):
            return torch.zeros(shape, dtype=dtype, device=device) if g is None else g.to(dtype).contiguous()

        ge = dense(grad_energy, (f,))
        gf = dense(grad_forces, (f, frames.num_atoms, 3)).reshape(-1, 3)
        gw = dense(grad_virial, (f, 3, 3))
        # the batch's bins and sorted positions are those of this forward's frames
        _launch(frames)
        grad_table = torch.zeros(ctx.table_shape[1], dtype=dtype, device=device)
        layout = b._layout(frames.schedule[1])
        _native.call("batch_table_vjp", _native.dtype_tag(dtype), _native.ctypes.byref(layout.native),
                     _native.ptr(ge), _native.ptr(gf), _native.ptr(gw), _native.ptr(frames.workspace),
                     _native.ctypes.c_int64(frames.workspace.numel()), _native.ptr(grad_table),
                     _native.stream(device))
        return grad_table.reshape(ctx.table_shape).to(ctx.table_dtype), None


def _launch(frames: TableFrames):
    frames.batch.launch(frames.schedule)
    b = frames.batch
    return b.energy, b.forces, b.virial


__all__ = ("TableFrames", "TabulatedEFW", "loss", "scales")
