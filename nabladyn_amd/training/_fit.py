"""Fitting a potential's parameters to labeled frames (``training/_fit.py:1-442``).

The frames are prepared once: moved to the model's device and dtype, with their pair stacks
(``pair_stack``). Each update takes a batch of frames, evaluates their channels and the loss
in chunks of ``batch_size`` frames, and writes the batch gradient into the parameters' ``.grad``
before the optimizer steps. For LJ and tabulated models the gradient comes from the parameter
VJP kernel (``training._loss``); no compiled program and no vmap are needed.
"""

import dataclasses
import math
from numbers import Integral, Real

import numpy as np
import torch
from torch import Tensor

from .. import _native, space
from ..neighbors import NeighborList
from ..potential import Potential
from . import _loss
from ._dataset import Dataset


def width(count: int) -> int:
    """``count`` rounded up to a multiple of 1/8 of its leading power of two
    (``training/_fit.py:190-197``): similar datasets share one stack width."""
    step = 1 << max(0, int(count).bit_length() - 4)
    return -(-int(count) // step) * step


def _frame_totals(wrapped: Tensor, cells: Tensor, cutoff: float, densest: int):
    """Pair totals [F] (device int32) of every frame from one zero-skin list sized at frame
    ``densest`` with twice its pairs, and the status words of that list. A list that
    overflows builds nothing more, so its totals from there on repeat the first overflowing
    frame's count."""
    frames = wrapped.shape[0]
    counter = NeighborList(cutoff, capacity_multiplier=2.0, full=False)
    counter.allocate(wrapped[densest], cells[densest])
    buffers = counter._live().buffers
    totals = torch.empty(frames, dtype=torch.int32, device=wrapped.device)
    for frame in range(frames):
        counter._rebuild_wrapped(wrapped[frame], cells[frame])
        totals[frame:frame + 1].copy_(buffers["pair_total"])
    return counter, totals, buffers["status"]


def pair_stack(positions: Tensor, cells: Tensor, cutoff: float) -> Tensor:
    """The pairs within ``cutoff`` of frames [F, N, 3] in cells [F, 3], int32 [F, 2, C]
    (``training/_fit.py:200-242``).

    Row f holds frame f's half list, then self-pairs (k mod N, k mod N) up to the common width
    C = ``width`` of the densest frame's capacity. A zero-skin list counts every frame's pairs;
    a list sized for the densest frame then builds each row. Syncs with the host a few times,
    not per frame, and again whenever a frame has over twice the pairs the count was sized for.

    Raises ``ValueError`` for a cell that does not fit the list or a non-finite coordinate.
    """
    frames, atoms = positions.shape[:2]
    cells = cells.detach().contiguous()
    wrapped = torch.stack([space.wrap(positions[f].detach(), cells[f])[0]
                           for f in range(frames)]).contiguous()
    densest = 0
    while True:
        counter, totals, status = _frame_totals(wrapped, cells, cutoff, densest)
        flags, densest = (int(v) for v in torch.stack(
            [status[0], totals.argmax().to(status.dtype)]).tolist())
        if flags & ~_native.STATUS_OVERFLOW:
            counter._check_status()
        if not flags & _native.STATUS_OVERFLOW:
            break
        # the first overflowing frame holds the largest count (argmax takes the first of
        # equals); count again with the list sized there
    sized = NeighborList(cutoff, capacity_multiplier=1.0, full=False)
    sized.allocate(wrapped[densest], cells[densest])
    capacity = sized.pair_capacity
    columns = width(capacity)
    fill = torch.arange(columns, dtype=torch.int32, device=positions.device) % atoms
    stack = fill.expand(frames, 2, columns).contiguous()
    index = sized.tensors().index
    for frame in range(frames):
        sized._rebuild_wrapped(wrapped[frame], cells[frame])
        stack[frame, :, :capacity].copy_(index)
    sized._check_status()
    return stack


def batches(frames: int, batch_size: int, updates: int, seed: int) -> list[Tensor]:
    """The frame indices of each update (``training/_fit.py:287-297``): every epoch a new
    shuffle from ``seed`` cut into batches of ``batch_size``, its last batch the remainder."""
    generator = torch.Generator().manual_seed(int(seed))
    out: list[Tensor] = []
    while len(out) < updates:
        out.extend(torch.randperm(frames, generator=generator).split(batch_size))
    return out[:updates]


def held_entries(model: Potential, trained: dict[str, Tensor]):
    """(parameter, mask of held entries, copy of its values) for each partly trained
    parameter (``training/_fit.py:300-307``)."""
    masks = model._trained_masks()
    return [(value, ~masks[name], value.detach().clone())
            for name, value in trained.items() if name in masks]


def optimizer_step(optimizer: torch.optim.Optimizer, trained: dict[str, Tensor],
                   gradient: dict[str, Tensor], held) -> None:
    """Write ``gradient`` into ``.grad`` and step (``training/_fit.py:310-329``). Held entries get
    zero gradient and their values back after the step, since weight decay moves them anyway."""
    optimizer.zero_grad()
    for name, value in trained.items():
        value.grad = gradient[name]
    for value, mask, _ in held:
        value.grad = value.grad.masked_fill(mask, 0)
    optimizer.step()
    with torch.no_grad():
        for value, mask, kept in held:
            value.copy_(torch.where(mask, kept, value))


@dataclasses.dataclass(frozen=True)
class Report:
    """The result of a fit (``training/_fit.py:174-187``).

    ``loss`` [updates]: each update's batch loss, before its step. ``scales``: the channel
    scales from the training labels. ``training`` and ``validation`` (None without a validation
    set): each weighted channel's RMS error (before, after) the fit, the energy per frame in
    kJ/mol, force components in kJ/mol/nm and virial components in kJ/mol.
    """

    loss: np.ndarray
    scales: dict[str, float]
    training: dict[str, tuple[float, float]]
    validation: dict[str, tuple[float, float]] | None


@dataclasses.dataclass
class _Prepared:
    """Frames on the model's device and dtype, wrapped into their cells, with one pair stack
    per list cutoff and, for the fused models, the full rows built from each stack."""

    positions: Tensor
    cells: Tensor
    stacks: tuple[Tensor, ...]
    labels: dict[str, Tensor]
    rows: tuple[_loss.FrameRows, ...] | None = None

    def chunk(self, frames: Tensor):
        """The stacks and rows of ``frames``."""
        stacks = tuple(stack[frames] for stack in self.stacks)
        rows = None if self.rows is None else tuple(r.select(frames) for r in self.rows)
        return stacks, rows


def _prepare(model: Potential, data: Dataset, channels: tuple[str, ...]) -> _Prepared:
    def moved(value: Tensor) -> Tensor:
        return value.to(device=model.device, dtype=model.dtype)

    positions, cells = moved(data.positions), moved(data.cells)
    positions = torch.stack([space.wrap(positions[f], cells[f])[0]
                             for f in range(positions.shape[0])]).contiguous()
    stacks = tuple(pair_stack(positions, cells, cutoff) for cutoff in model._list_cutoffs)
    rows = None
    if model._fused:
        rows = tuple(_loss.frame_rows(stack, model.num_atoms) for stack in stacks)
    return _Prepared(positions, cells, stacks,
                     {name: moved(data.labels[name]) for name in channels}, rows)


def _chunks(frames: Tensor, size: int):
    """The frame indices in chunks of at most ``size`` (no padding frames: nothing here is
    compiled for one chunk shape)."""
    yield from frames.split(size)


def batch_gradient(model: Potential, trained: dict[str, Tensor], prepared: _Prepared,
                   frames: Tensor, size: int, weights: dict[str, float],
                   channel_scales: dict[str, float]):
    """(gradient by name, loss, terms by name) of the mean loss over ``frames``, in chunks of
    ``size`` frames (``training/_fit.py:33-70``): each frame weighs 1/B, so the chunks' sums
    equal the batch's."""
    channels = tuple(prepared.labels)
    leaves = {name: value.detach().requires_grad_(True) for name, value in trained.items()}
    device = prepared.positions.device
    gradient = {name: torch.zeros_like(value) for name, value in trained.items()}
    total_loss, total_terms = None, None
    for chunk in _chunks(frames, size):
        chunk = chunk.to(device)
        coefficients = torch.full((len(chunk),), 1.0 / len(frames), dtype=model.dtype,
                                  device=device)
        stacks, rows = prepared.chunk(chunk)
        predicted = _loss._predict(model, leaves, prepared.positions[chunk],
                                   prepared.cells[chunk], stacks, channels, wrapped=True,
                                   rows=rows)
        labels = {name: prepared.labels[name][chunk] for name in channels}
        value, terms = _loss.loss(predicted, labels, weights, channel_scales, coefficients)
        grads = torch.autograd.grad(value, tuple(leaves.values()), allow_unused=True)
        for name, g in zip(leaves, grads):
            if g is not None:
                gradient[name] += g
        value, terms = value.detach(), {k: v.detach() for k, v in terms.items()}
        if total_loss is None:
            total_loss, total_terms = value, terms
        else:
            total_loss = total_loss + value
            total_terms = {k: total_terms[k] + v for k, v in terms.items()}
    return gradient, total_loss, total_terms


def _errors(model: Potential, prepared: _Prepared, size: int) -> dict[str, float]:
    """Each channel's RMS error over a prepared set at the model's own parameters."""
    channels = tuple(prepared.labels)
    squares = {name: 0.0 for name in channels}
    frames = torch.arange(prepared.positions.shape[0])
    with torch.no_grad():
        for chunk in _chunks(frames, size):
            chunk = chunk.to(prepared.positions.device)
            stacks, rows = prepared.chunk(chunk)
            predicted = _loss._predict(model, {}, prepared.positions[chunk],
                                       prepared.cells[chunk], stacks, channels, wrapped=True,
                                       rows=rows)
            for name in channels:
                error = predicted[name] - prepared.labels[name][chunk]
                squares[name] = squares[name] + error.double().square().sum()
    return {name: math.sqrt(float(squares[name]) / prepared.labels[name].numel())
            for name in channels}


def _check_arguments(model, weights, counts, frames) -> None:
    """The argument checks of ``training/_fit.py:103-171``."""
    if not isinstance(model, Potential):
        raise TypeError("model must be a Potential")
    if not weights:
        raise ValueError("give at least one channel's weight: energy, forces or virial")
    for name, value in weights.items():
        if (isinstance(value, bool) or not isinstance(value, Real) or not math.isfinite(value)
                or value <= 0):
            raise ValueError(f"the {name} weight must be a finite positive number")
    for name, value in counts.items():
        if isinstance(value, bool) or not isinstance(value, Integral) or value < 0:
            raise ValueError(f"{name} must be a nonnegative integer")
    if counts["updates"] == 0 or counts["echo_every"] == 0 or not 0 < counts["batch_size"] <= frames:
        raise ValueError("updates and echo_every must be positive, and batch_size between 1 and "
                         "the training set's size")


def _check_sets(model: Potential, sets: dict[str, Dataset], channels) -> None:
    for label, data in sets.items():
        atoms = data.positions.shape[1]
        if atoms != model.num_atoms:
            raise ValueError(f"the {label} frames hold {atoms} atoms, the model {model.num_atoms}")
        for name in channels:
            if name not in data.labels:
                raise ValueError(f"the {label} frames have no {name} labels")


def fit(model: Potential, train: Dataset, optimizer: torch.optim.Optimizer, *,
        updates: Integral, batch_size: Integral, energy: Real | None = None,
        forces: Real | None = None, virial: Real | None = None,
        validation: Dataset | None = None, seed: Integral = 0,
        scheduler: torch.optim.lr_scheduler.LRScheduler | None = None,
        echo_every: Integral | None = None) -> Report:
    r"""Fit the model's trained parameters to the training labels; return a ``Report``
    (``training/_fit.py:332-439``).

    Trained parameters are those that require gradients. A channel enters the loss when its
    weight w is given (``energy``, ``forces``, ``virial``), as w mean[((X - X*) / s)^2] with
    the scale s from the training labels. Each of the ``updates`` takes the next batch of
    ``batch_size`` frames (each epoch a new shuffle from ``seed``, its last batch the
    remainder), writes the gradient of the batch's mean loss into ``.grad`` and calls
    ``optimizer.step()``. Rows the model does not train keep their values. Parameters change
    in place; another call continues from them and the optimizer. ``scheduler`` steps after
    each update (not one that needs a metric). The report gives the RMS errors before and
    after, on the training frames and on ``validation``.

    With ``echo_every`` = k, prints a header, then every k updates the update number, the mean
    loss and the mean unweighted terms over those k updates. Syncs with the host.

    Raises ``TypeError`` for a model that is not a Potential, and ``ValueError`` for weights
    or counts out of range, a model that trains no parameter, or frames with another atom
    count or without the weighted labels.
    """
    weights = {name: value for name, value in (("energy", energy), ("forces", forces),
                                               ("virial", virial)) if value is not None}
    counts = {"updates": updates, "batch_size": batch_size, "seed": seed,
              "echo_every": 1 if echo_every is None else echo_every}
    _check_arguments(model, weights, counts, len(train))
    trained = {name: value for name, value in model.named_parameters() if value.requires_grad}
    if not trained:
        raise ValueError("the model has no trainable parameter; set trainable when building "
                         "the potential")
    held = held_entries(model, trained)
    channels = tuple(name for name in _loss.CHANNELS if name in weights)
    sets = {"training": train}
    if validation is not None:
        sets["validation"] = validation
    _check_sets(model, sets, channels)
    prepared = {label: _prepare(model, data, channels) for label, data in sets.items()}
    channel_scales = _loss.scales(**prepared["training"].labels)
    before = {label: _errors(model, data, batch_size) for label, data in prepared.items()}
    losses = torch.empty(updates, dtype=torch.float64, device=model.device)
    terms = torch.empty((updates, len(channels)), dtype=torch.float64, device=model.device)
    if echo_every is not None:
        print("update loss " + " ".join(channels), flush=True)
    for update, frames in enumerate(batches(len(train), batch_size, updates, seed)):
        gradient, value, parts = batch_gradient(model, trained, prepared["training"], frames,
                                                batch_size, weights, channel_scales)
        optimizer_step(optimizer, trained, gradient, held)
        if scheduler is not None:
            scheduler.step()
        losses[update] = value
        terms[update] = torch.stack([parts[name] for name in channels])
        if echo_every is not None and (update + 1) % echo_every == 0:
            window = slice(update + 1 - echo_every, update + 1)
            means = torch.cat([losses[window].mean().reshape(1), terms[window].mean(dim=0)])
            print(f"{update + 1} " + " ".join(f"{v:.6g}" for v in means.tolist()), flush=True)
    after = {label: _errors(model, data, batch_size) for label, data in prepared.items()}
    errors = {label: {name: (before[label][name], after[label][name]) for name in channels}
              for label in prepared}
    return Report(loss=losses.cpu().numpy(), scales=channel_scales,
                  training=errors["training"], validation=errors.get("validation"))


__all__ = ("Report", "batch_gradient", "batches", "fit", "held_entries", "optimizer_step",
           "pair_stack", "width")
