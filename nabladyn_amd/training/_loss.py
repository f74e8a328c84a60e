"""The supervised loss and the channels it reads (``training/_loss.py:1-128``).

``predict`` gives the energy, forces and virial of B frames, differentiable in the parameters it
is handed. LJ and tabulated models (and sums of them) take one launch of the fused pair kernel
per part for all the frames; the backward of that launch is the parameter VJP kernel, which
contracts the upstream gradients of E, F and W with each pair's parameter derivatives. That
is the mixed second derivative the reference forms with nested ``torch.func.grad``
(``training/_loss.py:94-102``), computed directly. The frames of a batch go through each kernel
together (``ndn_pair_efw_frames``, ``ndn_pair_parameter_vjp_frames``), where the reference
vmaps. Any other model is differentiated twice by autograd, as the reference does.
"""

from collections.abc import Mapping
from typing import NamedTuple

import torch
from torch import Tensor, nn

from .. import space
from ..neighbors import NeighborPairs
from ..potential import Potential, Sum
from ..potential._energy import evaluate_frames, evaluate_frames_full, parameter_vjp_frames

CHANNELS = ("energy", "forces", "virial")


def scales(energy: Tensor | None = None, forces: Tensor | None = None,
           virial: Tensor | None = None) -> dict[str, float]:
    """Each given channel's scale from its training labels (``training/_loss.py:26-44``).

    Energy [F] and virial [F, 3, 3]: the RMS about the mean over frames; forces [F, N, 3]: the
    plain RMS. Labels constant over frames fall back to the plain RMS, all-zero labels to 1.
    Syncs with the host.
    """
    out = {}
    for name, labels, centred in (("energy", energy, True), ("forces", forces, False),
                                  ("virial", virial, True)):
        if labels is None:
            continue
        plain = float(labels.square().mean().sqrt())
        spread = float((labels - labels.mean(dim=0)).square().mean().sqrt()) if centred else plain
        out[name] = spread or plain or 1.0
    return out


def loss(predicted: Mapping[str, Tensor], labels: Mapping[str, Tensor],
         weights: Mapping[str, float | Tensor], scales: Mapping[str, float | Tensor],
         coefficients: Tensor | None = None) -> tuple[Tensor, dict[str, Tensor]]:
    """A batch's loss, sum_X w_X term_X, and the terms by name (``training/_loss.py:105-125``).

    term_X sums over the B frames each frame's mean of ((X - X*) / s_X)^2 over its components,
    times the frame's coefficient [B] (1/B without ``coefficients``).
    """
    terms = {}
    for name, value in predicted.items():
        scaled = ((value - labels[name]) / scales[name]).square()
        per_frame = scaled.reshape(scaled.shape[0], -1).mean(dim=1)
        terms[name] = per_frame.mean() if coefficients is None else (coefficients * per_frame).sum()
    total = None
    for name, term in terms.items():
        total = weights[name] * term if total is None else total + weights[name] * term
    return total, terms


class FrameRows(NamedTuple):
    """Full rows of frames built from their pair stacks: row r of frame f (its atom r) lists every
    pair of the stack with r, in both orientations, 64 rows per slice, column-major within a
    slice: the k-th neighbor of row r at index[f, slices[f, r // 64] + r % 64 + 64 k] for
    k < counts[f, r]. All int32."""

    counts: Tensor  # [F, N]
    slices: Tensor  # [F, ceil(N / 64) + 1]
    index: Tensor   # [F, capacity]

    def select(self, frames: Tensor) -> "FrameRows":
        return FrameRows(self.counts[frames], self.slices[frames], self.index[frames])


def frame_rows(stack: Tensor, num_atoms: int) -> FrameRows:
    """The full rows of the frames of a pair stack [F, 2, C] (self-pairs are padding). Each row
    keeps the stack's order: its pairs as first atom, then as second. Built once per prepared set
    (the frames do not move), so the fused forward sums each atom's force in registers."""
    frames, _, columns = stack.shape
    device = stack.device
    first, second = stack[:, 0].long(), stack[:, 1].long()
    listed = first != second
    frame = torch.arange(frames, device=device)[:, None].expand(frames, columns)
    owner = torch.cat([frame[listed], frame[listed]])
    row = torch.cat([first[listed], second[listed]])
    neighbor = torch.cat([second[listed], first[listed]])
    key = owner * num_atoms + row
    order = torch.argsort(key, stable=True)
    key, neighbor = key[order], neighbor[order]
    counts = torch.bincount(key, minlength=frames * num_atoms)
    starts = torch.cumsum(counts, 0) - counts
    rank = torch.arange(key.numel(), device=device) - starts[key]
    per_frame = -(-num_atoms // 64)
    padded = torch.zeros(frames, per_frame * 64, dtype=counts.dtype, device=device)
    padded[:, :num_atoms] = counts.view(frames, num_atoms)
    width = padded.view(frames, per_frame, 64).amax(dim=2)
    slices = torch.zeros(frames, per_frame + 1, dtype=torch.int64, device=device)
    slices[:, 1:] = torch.cumsum(64 * width, dim=1)
    capacity = max(int(slices[:, -1].max()), 1)
    index = torch.zeros(frames, capacity, dtype=torch.int32, device=device)
    f, r = key // num_atoms, key % num_atoms
    index[f, slices[f, r // 64] + r % 64 + 64 * rank] = neighbor.to(torch.int32)
    return FrameRows(counts.view(frames, num_atoms).to(torch.int32).contiguous(),
                     slices.to(torch.int32).contiguous(), index)


class _FusedFrames(torch.autograd.Function):
    """(E [B], F [B, N, 3], W [B, 3, 3]) of an LJ or tabulated model on B frames (wrapped
    positions, pair stacks [B, 2, C]), differentiable in the model's two parameter tensors (LJ:
    sigma and epsilon; a table: the table and None). Forward: ``ndn_pair_efw_full_frames`` over
    the frames' full rows when they are given, else ``ndn_pair_efw_frames`` over the stacks, one
    launch for every frame; backward: ``ndn_pair_parameter_vjp_frames`` over the stacks, one
    launch. The positions and the cells are data here and get no gradient."""

    @staticmethod
    def forward(positions, cells, stack, model, first, second, rows):
        native = model._native()
        if rows is not None:
            e, f, w = evaluate_frames_full(model, positions, cells, rows, native)
        else:
            e, f, w = evaluate_frames(model, positions, cells, stack, native)
        return e, f, w, native

    @staticmethod
    def setup_context(ctx, inputs, output):
        positions, cells, stack, model, first, second, rows = inputs
        ctx.set_materialize_grads(False)
        ctx.model, ctx.native = model, output[3]
        ctx.save_for_backward(positions, cells, stack)

    @staticmethod
    def backward(ctx, grad_energy, grad_forces, grad_virial, _):
        positions, cells, stack = ctx.saved_tensors
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            raise RuntimeError("training channels give no gradient in the positions or the cells")
        wanted = ctx.needs_input_grad[4] or ctx.needs_input_grad[5]
        if not wanted or (grad_energy is None and grad_forces is None and grad_virial is None):
            return None, None, None, None, None, None, None
        g = parameter_vjp_frames(ctx.model, positions, cells, stack, grad_energy, grad_forces,
                                 grad_virial, ctx.native)
        first, second = ctx.model._split_parameter_gradient(g.to(positions.dtype))
        return (None, None, None, None, first if ctx.needs_input_grad[4] else None,
                second if ctx.needs_input_grad[5] else None, None)


def _list_for(lists: dict, cutoff: float):
    """The entry of the smallest list cutoff that reaches ``cutoff`` (``list_for``'s rule)."""
    return lists[min(c for c in lists if c >= cutoff)]


def _fused_frames(model, wrapped: Tensor, cells: Tensor, stacks: dict[float, Tensor],
                  rows: dict[float, FrameRows] | None):
    """(E, F, W) of a fused model (LJ, a table, or a sum of fused models) on B frames, each
    part on the shortest list that reaches its cutoff, added in part order."""
    if isinstance(model, Sum):
        parts = [_fused_frames(part, wrapped, cells, stacks, rows) for part in model.parts]
        return tuple(sum(values[1:], values[0]) for values in zip(*parts))
    stack = _list_for(stacks, model.cutoff)
    if stack.dtype != torch.int32 or not stack.is_contiguous():
        stack = stack.to(torch.int32).contiguous()
    own_rows = None if rows is None else _list_for(rows, model.cutoff)
    first, second = model._differentiable_parameters()
    e, f, w, _ = _FusedFrames.apply(wrapped, cells, stack, model, first, second, own_rows)
    return e, f, w


def _autograd_frame(model, positions: Tensor, cell: Tensor, pairs: NeighborPairs):
    """(E, F, W) of any differentiable model by autograd with a graph kept (the reference's
    route, ``quantity.py:21-117``), so the loss differentiates F and W again."""
    x = positions.detach().requires_grad_(True)
    strain = torch.zeros((3, 3), dtype=positions.dtype, device=positions.device,
                         requires_grad=True)
    e = model(x, cell, pairs, strain=strain)
    gx, gs = torch.autograd.grad(e, (x, strain), create_graph=True)
    return e, -gx, -gs


class _Channels(nn.Module):
    """The model under ``functional_call``: its parameters replaced by the ones handed in."""

    def __init__(self, model: Potential, wrapped: bool):
        super().__init__()
        self.model = model
        self.wrapped = wrapped

    def forward(self, positions: Tensor, cells: Tensor, stacks: tuple[Tensor, ...],
                rows: tuple[FrameRows, ...] | None = None):
        model = self.model
        cutoffs = model._list_cutoffs
        if model._fused:
            x = positions
            if not self.wrapped:
                x = torch.stack([space.wrap(positions[f], cells[f])[0]
                                 for f in range(positions.shape[0])])
            return _fused_frames(model, x.contiguous(), cells.contiguous(),
                                 dict(zip(cutoffs, stacks)),
                                 None if rows is None else dict(zip(cutoffs, rows)))
        out = []
        for frame in range(positions.shape[0]):
            lists = [NeighborPairs(stack[frame], model.num_atoms, cutoff)
                     for stack, cutoff in zip(stacks, cutoffs)]
            pairs = NeighborPairs(lists[-1].index, model.num_atoms, cutoffs[-1],
                                  tuple(lists[:-1]))
            out.append(_autograd_frame(model, positions[frame], cells[frame], pairs))
        return tuple(torch.stack([values[k] for values in out]) for k in range(3))


def predict(model: Potential, parameters: Mapping[str, Tensor], positions: Tensor,
            cells: Tensor, stacks: tuple[Tensor, ...],
            channels: tuple[str, ...] = CHANNELS) -> dict[str, Tensor]:
    """The channels of B frames by name (``training/_loss.py:71-102``).

    Takes positions [B, N, 3], cells [B, 3] and pair indices [B, 2, C_k], one per cutoff of the
    model's ``_list_cutoffs`` (ascending; the last is the list, the others its inner lists),
    padded with self-pairs that count nothing; gives energy [B], forces [B, N, 3] and virial
    [B, 3, 3]. ``parameters`` replaces the named model parameters, so the channels are
    differentiable in them. ``channels`` is a nonempty subsequence of ``CHANNELS``.
    """
    return _predict(model, parameters, positions, cells, stacks, channels, wrapped=False)


def _predict(model, parameters, positions, cells, stacks, channels, *, wrapped: bool,
             rows=None):
    """``predict``; ``wrapped`` says the positions already lie in [0, L) (wrapped by the
    library's wrap, as a prepared set's are), so the fused path skips wrapping them again;
    ``rows`` (one ``FrameRows`` per stack) lets the fused forward run over full rows."""
    wrapper = _Channels(model, wrapped)
    named = {f"model.{name}": value for name, value in parameters.items()}
    values = torch.func.functional_call(wrapper, named,
                                        (positions, cells, tuple(stacks), rows))
    return {name: values[CHANNELS.index(name)] for name in channels}


__all__ = ("CHANNELS", "loss", "predict", "scales")
