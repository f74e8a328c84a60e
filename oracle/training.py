"""ORACLE — test infrastructure, not product: the fitting gradient restated on the CPU.

The reference's fit differentiates the loss of predicted energies, forces and virials with
respect to the model parameters (``training/_loss.py:71-125``, ``training/_fit.py:245-271``):
F = -dE/dx and W = -dE/dstrain are formed with a graph kept, then the scaled MSE loss is
differentiated again. Here that is eager torch autograd with ``create_graph=True`` over the
oracle's torch energies (``oracle.potentials.t_lj_energy``, ``t_table_energy``), in any dtype.

Parity status: pinned in tests/test_oracle_training.py by central differences of the loss in
fp64 (LJ sigma and epsilon of a two-type model, and a two-type table), with a planted mistake
(forces taken without their graph) that the differences catch.
"""

import torch

CHANNELS = ("energy", "forces", "virial")


def frame_channels(energy, positions, cell, index):
    """(E, F, W) of ``energy(positions, strain, cell, index)`` with the graph kept."""
    x = positions.detach().clone().requires_grad_(True)
    strain = torch.zeros((3, 3), dtype=x.dtype, requires_grad=True)
    e = energy(x, strain, cell, index)
    gx, gs = torch.autograd.grad(e, (x, strain), create_graph=True)
    return e, -gx, -gs


def loss(predicted, labels, weights, scales):
    """sum_X w_X mean over frames of each frame's mean ((X - X*) / s_X)^2
    (``training/_loss.py:105-125`` with every frame weighing 1/B)."""
    total = 0.0
    for name, value in predicted.items():
        squares = ((value - labels[name]) / scales[name]).square()
        total = total + weights[name] * squares.reshape(squares.shape[0], -1).mean(dim=1).mean()
    return total


def loss_and_gradient(energy, parameters, frames, labels, weights, scales):
    """(loss, d loss / d parameters) over ``frames`` [(positions, cell, index)], ``energy``
    taking (parameters, positions, strain, cell, index)."""
    leaves = [p.detach().clone().requires_grad_(True) for p in parameters]
    channels = tuple(name for name in CHANNELS if name in weights)
    outputs = {name: [] for name in channels}
    for positions, cell, index in frames:
        values = frame_channels(lambda x, s, c, i: energy(leaves, x, s, c, i), positions, cell,
                                index)
        for name, value in zip(CHANNELS, values):
            if name in outputs:
                outputs[name].append(value)
    predicted = {name: torch.stack(values) for name, values in outputs.items()}
    value = loss(predicted, {k: labels[k] for k in channels}, weights, scales)
    if not value.requires_grad:  # nothing in the loss depends on the parameters
        return value.detach(), [torch.zeros_like(p) for p in leaves]
    grads = torch.autograd.grad(value, leaves, allow_unused=True)
    return value.detach(), [torch.zeros_like(p) if g is None else g
                            for p, g in zip(leaves, grads)]


__all__ = ("frame_channels", "loss", "loss_and_gradient")
