"""Pins the oracle's fitting gradient (``oracle/training.py``), the autograd restatement of the
reference's double backward (``training/_loss.py:71-125``), by central differences of the loss in
fp64 on the CPU:

- a two-type LJ model: E, F and W are linear in epsilon, so the loss is quadratic in it and the
  differences are exact up to rounding; in sigma they converge as h^2 (h = 1e-5);
- a two-type table: the loss is quadratic in the knots; the last knot's gradient is 0;
- a planted mistake (forces taken without their graph, so the force term adds nothing to the
  gradient) is caught.
"""

import numpy as np
import torch

from nabladyn_amd import training
from oracle import neighbors as on
from oracle import potentials as op
from oracle import training as ot

CUTOFF = 1.0
WEIGHTS = {"energy": 1.0, "forces": 1.0, "virial": 0.5}


def _frames(seed=3, count=2):
    rng = np.random.default_rng(seed)
    types = torch.tensor(rng.integers(0, 2, 216))
    frames = []
    for k in range(count):
        cell = np.full(3, 2.4)
        x = on.jittered_lattice(6, 2.4, 0.03, seed=seed + k)
        xw = on.wrap(x, cell)[0]
        index = torch.tensor(on.brute_pairs(xw, cell, CUTOFF).T).long()
        frames.append((torch.tensor(xw), torch.tensor(cell), index))
    labels = {"energy": torch.tensor(rng.normal(-300, 20, count)),
              "forces": torch.tensor(rng.normal(0, 50, (count, 216, 3))),
              "virial": torch.tensor(rng.normal(0, 40, (count, 3, 3)))}
    return frames, types, labels


def _lj(types):
    def energy(params, x, strain, cell, index):
        sigma, epsilon = params
        return op.t_lj_energy(x, strain, cell, index, types, sigma, epsilon, CUTOFF)
    return energy


def _table(types, knots=30):
    start = torch.full((3,), 0.3, dtype=torch.float64)
    spacing = (CUTOFF - start) / (knots - 1)

    def energy(params, x, strain, cell, index):
        (table,) = params
        return op.t_table_energy(x, strain, cell, index, types, table, start, spacing, CUTOFF)
    return energy


def _loss_at(energy, params, frames, labels, scales, weights):
    value, _ = ot.loss_and_gradient(energy, params, frames, labels, weights, scales)
    return float(value)


def _central(energy, params, which, entry, h, frames, labels, scales, weights=WEIGHTS):
    up = [p.clone() for p in params]
    down = [p.clone() for p in params]
    up[which].view(-1)[entry] += h
    down[which].view(-1)[entry] -= h
    return (_loss_at(energy, up, frames, labels, scales, weights)
            - _loss_at(energy, down, frames, labels, scales, weights)) / (2 * h)


def test_lj_gradient_equals_central_differences():
    frames, types, labels = _frames()
    scales = training.scales(**labels)
    params = [torch.tensor([0.33, 0.35, 0.37], dtype=torch.float64),
              torch.tensor([0.9, 1.05, 1.2], dtype=torch.float64)]
    energy = _lj(types)
    _, (g_sigma, g_epsilon) = ot.loss_and_gradient(energy, params, frames, labels, WEIGHTS, scales)
    for row in range(3):
        fd = _central(energy, params, 1, row, 0.05, frames, labels, scales)
        assert abs(fd - float(g_epsilon[row])) <= 1e-9 * float(g_epsilon.abs().max()), row
        fd = _central(energy, params, 0, row, 1e-5, frames, labels, scales)
        assert abs(fd - float(g_sigma[row])) <= 1e-6 * float(g_sigma.abs().max()), row


def test_table_gradient_equals_central_differences():
    frames, types, labels = _frames(seed=5)
    scales = training.scales(**labels)
    rng = np.random.default_rng(6)
    r = 0.3 + 0.7 / 29 * np.arange(30)
    base = np.clip(4.0 * ((0.36 / r) ** 12 - (0.36 / r) ** 6), -40, 40)
    table = torch.tensor(np.stack([base + rng.normal(0, 0.1, 30) for _ in range(3)]))
    table[:, -1] = 0.0
    energy = _table(types)
    _, (grad,) = ot.loss_and_gradient(energy, [table], frames, labels, WEIGHTS, scales)
    assert torch.all(grad[:, -1] == 0)
    scale = float(grad.abs().max())
    for row, knot in ((0, 0), (0, 3), (1, 10), (2, 17), (2, 28)):
        fd = _central(energy, [table], 0, row * 30 + knot, 0.5, frames, labels, scales)
        assert abs(fd - float(grad[row, knot])) <= 1e-9 * scale, (row, knot)


def test_forces_without_their_graph_are_caught():
    frames, types, labels = _frames(seed=9)
    scales = training.scales(**labels)
    params = [torch.tensor([0.33, 0.35, 0.37], dtype=torch.float64),
              torch.tensor([0.9, 1.05, 1.2], dtype=torch.float64)]
    energy = _lj(types)
    forces_only = {"forces": 1.0}

    def detached_channels(energy_fn, positions, cell, index):
        e, f, w = original(energy_fn, positions, cell, index)
        return e, f.detach(), w

    original = ot.frame_channels
    try:
        ot.frame_channels = detached_channels
        _, (_, broken) = ot.loss_and_gradient(energy, params, frames, labels, forces_only,
                                              scales)
    finally:
        ot.frame_channels = original
    _, (_, right) = ot.loss_and_gradient(energy, params, frames, labels, forces_only, scales)
    fd = _central(energy, params, 1, 1, 0.05, frames, labels, scales, forces_only)
    assert abs(fd - float(right[1])) <= 1e-9 * abs(fd)
    assert abs(fd - float(broken[1])) > 0.5 * abs(fd)
