"""Pins the gradient oracle of table fitting: the autograd restatement of the reference's double
backward (F = -dE/dx, W = -dE/dstrain with create_graph, the loss of training/_loss.py:105-125,
then d/dtable) equals fp64 central differences of the closed-form loss. E, F and W are linear in
the knots, so the loss is quadratic in them and the differences are exact up to rounding."""

import numpy as np
import torch

from nabladyn_amd import fit, workloads
from oracle import neighbors as on
from oracle import potentials as op

KNOTS, START, CUTOFF = 40, 0.3, 1.2
WEIGHTS = {"energy": 1.0, "forces": 1.0, "virial": 1.0}


def _frames():
    xs, cells, hidden = workloads.cg_frames(frames=2, beads=216, seed=7)
    rng = np.random.default_rng(8)
    r = START + (CUTOFF - START) / (KNOTS - 1) * np.arange(KNOTS)
    s6 = (0.45 / r) ** 6
    table = np.clip(12.0 * (s6 * s6 - s6), -50, 50) + rng.normal(0, 0.05, KNOTS)
    table[-1] = 0.0
    labels = {"energy": rng.normal(size=2), "forces": rng.normal(size=xs.shape) * 5,
              "virial": rng.normal(size=(2, 3, 3))}
    pairs = [on.brute_pairs(on.wrap(x, c)[0], c, CUTOFF).T for x, c in zip(xs, cells)]
    return xs, cells, pairs, table, labels


def _predict(xs, cells, pairs, table):
    spacing = np.array([(CUTOFF - START) / (KNOTS - 1)])
    out = [op.table_efw(on.wrap(x, c)[0], c, p, np.zeros(len(x), np.int64), table[None],
                        np.array([START]), CUTOFF, spacing=spacing)
           for x, c, p in zip(xs, cells, pairs)]
    return {"energy": np.array([o[0] for o in out]), "forces": np.stack([o[1] for o in out]),
            "virial": np.stack([o[2] for o in out])}


def _loss(pred, labels, scales):
    return sum(WEIGHTS[k] * (((v - labels[k]) / scales[k]) ** 2).reshape(len(v), -1).mean(1).mean()
               for k, v in pred.items())


def test_autograd_gradient_equals_central_differences():
    xs, cells, pairs, table, labels = _frames()
    scales = fit.scales(**{k: torch.tensor(v) for k, v in labels.items()})
    t = torch.tensor(table, requires_grad=True)
    spacing = torch.tensor([(CUTOFF - START) / (KNOTS - 1)], dtype=torch.float64)
    pred = {"energy": [], "forces": [], "virial": []}
    for x, c, p in zip(xs, cells, pairs):
        pos = torch.tensor(on.wrap(x, c)[0], requires_grad=True)
        strain = torch.zeros((3, 3), dtype=torch.float64, requires_grad=True)
        e = op.t_table_energy(pos, strain, torch.tensor(c), torch.tensor(p).long(),
                              torch.zeros(len(x), dtype=torch.int64), t[None],
                              torch.tensor([START], dtype=torch.float64), spacing, CUTOFF)
        gx, gs = torch.autograd.grad(e, (pos, strain), create_graph=True)
        pred["energy"].append(e)
        pred["forces"].append(-gx)
        pred["virial"].append(-gs)
    value, _ = fit.loss({k: torch.stack(v) for k, v in pred.items()},
                        {k: torch.tensor(v) for k, v in labels.items()}, WEIGHTS, scales)
    (grad,) = torch.autograd.grad(value, t)
    grad = grad.numpy()
    assert grad[-1] == 0.0
    h = 0.5
    for k in range(KNOTS - 1):
        up, down = table.copy(), table.copy()
        up[k] += h
        down[k] -= h
        fd = (_loss(_predict(xs, cells, pairs, up), labels, scales) -
              _loss(_predict(xs, cells, pairs, down), labels, scales)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-10 * np.abs(grad).max(), k
    # a planted mistake — the gradient of a loss without the virial term — is caught
    partial = dict(WEIGHTS, virial=0.0)
    fd_without = (_loss_weighted(_predict(xs, cells, pairs, table + _unit(3, h)), labels, scales, partial) -
                  _loss_weighted(_predict(xs, cells, pairs, table - _unit(3, h)), labels, scales, partial)) / (2 * h)
    assert abs(fd_without - grad[3]) > 1e-6 * np.abs(grad).max()


def _unit(k, h):
    e = np.zeros(KNOTS)
    e[k] = h
    return e


def _loss_weighted(pred, labels, scales, weights):
    return sum(weights[k] * (((v - labels[k]) / scales[k]) ** 2).reshape(len(v), -1).mean(1).mean()
               for k, v in pred.items())
