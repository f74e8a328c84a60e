"""The hand-written backward of the tabulated potential (``fit.TabulatedEFW``): predictions and
dLoss/dknots on small frames shaped like the force-matching workload (beads at 8 nm^-3, one
type, a 200-knot table on [0.3, 1.2] nm, targets from a hidden LJ plus noise on the forces).

The gradient is checked two ways, both in fp64 on the same (rounded) positions:
  * central differences of the oracle's loss — E, F and W are linear in the knots, so the loss is
    quadratic in them and a central difference is exact up to rounding, whatever the step;
  * the autograd restatement of the reference's double backward: F = -dE/dx and W = -dE/dstrain
    built with create_graph, the loss of ``training/_loss.py:105-125``, then d/dtable.
Tolerances (max |dg| / max |g_ref| over the knots): fp64 1e-9; fp32 max(1e-4, 1.5 x the
distance of the reference's own route evaluated in fp32 from fp64), which is 1.8e-4 on these
frames (per-pair terms in fp32 here, sums in fp64). The last knot's gradient is exactly 0 (``_tabulated/_potential.py:143-144``).
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import fit, potential, workloads
from oracle import neighbors as on
from oracle import potentials as op

pytestmark = pytest.mark.gpu

KNOTS, START, CUTOFF = 200, 0.3, 1.2
WEIGHTS = {"energy": 1.0, "forces": 1.0, "virial": 1.0}


def _pairs(xw, cell):
    """Unordered pairs within the cutoff [2, P] from the oracle's C cell list (pinned by brute
    force in the oracle's own tests)."""
    b = on.build(xw, cell, CUTOFF)
    return b.pairs[:, :b.pair_total]


def _problem(frames=4, beads=512, seed=4):
    xs, cells, hidden = workloads.cg_frames(frames=frames, beads=beads, seed=seed)
    rng = np.random.default_rng(seed + 1)
    spacing = (CUTOFF - START) / (KNOTS - 1)
    r = START + spacing * np.arange(KNOTS)
    # a starting table: a softer LJ than the hidden one, plus noise, last knot 0
    s6 = (0.45 / r) ** 6
    table = 4 * 3.0 * (s6 * s6 - s6) + rng.normal(0, 0.05, KNOTS)
    table = np.clip(table, -50, 50)
    table[-1] = 0.0
    labels = {"energy": [], "forces": [], "virial": []}
    for x, cell in zip(xs, cells):
        xw = on.wrap(x, cell)[0]
        pairs = _pairs(xw, cell)
        e, f, w = op.lj_efw(xw, cell, pairs, np.zeros(len(xw), np.int64), [hidden["sigma"]],
                            [hidden["epsilon"]], CUTOFF)
        labels["energy"].append(e)
        labels["forces"].append(f + hidden["rng"].normal(0, hidden["force_noise"], f.shape))
        labels["virial"].append(w)
    labels = {k: np.stack(v) for k, v in labels.items()}
    return xs, cells, table, labels


def _oracle_predict(xs, cells, table):
    """fp64 E [F], F [F, N, 3], W [F, 3, 3] of the table on each frame (closed form)."""
    spacing = np.array([(CUTOFF - START) / (KNOTS - 1)])
    out = {"energy": [], "forces": [], "virial": []}
    for x, cell in zip(xs, cells):
        xw = on.wrap(x, cell)[0]
        pairs = _pairs(xw, cell)
        e, f, w = op.table_efw(xw, cell, pairs, np.zeros(len(xw), np.int64), table[None],
                               np.array([START]), CUTOFF, spacing=spacing)
        out["energy"].append(e)
        out["forces"].append(f)
        out["virial"].append(w)
    return {k: np.stack(v) for k, v in out.items()}


def _np_loss(pred, labels, scales):
    total = 0.0
    for name, value in pred.items():
        sq = ((value - labels[name]) / scales[name]) ** 2
        total += WEIGHTS[name] * sq.reshape(sq.shape[0], -1).mean(axis=1).mean()
    return total


def _autograd_gradient(xs, cells, table, labels, scales, dtype=torch.float64):
    """The reference's route restated in torch (fp64 unless asked): nested autograd through
    E(x, strain), the loss of ``training/_loss.py:105-125``, then d/dtable. The loss is a mean
    over frames of per-frame terms, so its gradient is accumulated frame by frame."""
    t = torch.tensor(table, dtype=dtype, requires_grad=True)
    spacing = torch.tensor([(CUTOFF - START) / (KNOTS - 1)], dtype=dtype)
    start = torch.tensor([START], dtype=dtype)
    count = len(xs)
    total = torch.zeros_like(t)
    for f, (x, cell) in enumerate(zip(xs, cells)):
        xw = on.wrap(x, cell)[0]
        pairs = torch.tensor(_pairs(xw, cell)).long()
        pos = torch.tensor(xw, dtype=dtype, requires_grad=True)
        strain = torch.zeros((3, 3), dtype=dtype, requires_grad=True)
        types = torch.zeros(len(xw), dtype=torch.int64)
        e = op.t_table_energy(pos, strain, torch.tensor(cell, dtype=dtype), pairs, types,
                              t[None], start, spacing, CUTOFF)
        gx, gs = torch.autograd.grad(e, (pos, strain), create_graph=True)
        pred = {"energy": e, "forces": -gx, "virial": -gs}
        term = 0.0
        for name, value in pred.items():
            target = torch.tensor(labels[name][f], dtype=dtype)
            term = term + WEIGHTS[name] * (((value - target) / scales[name]) ** 2).mean()
        (grad,) = torch.autograd.grad(term / count, t)
        total += grad
    return total.double().numpy()


def _frames(xs, cells, table, dtype, device):
    types = torch.zeros(xs.shape[1], dtype=torch.int64, device=device)
    model = potential.tabulated(atom_types=types,
                                table=torch.tensor(table[None], dtype=dtype, device=device),
                                start=torch.tensor([START], dtype=dtype, device=device),
                                cutoff=CUTOFF)
    return fit.TableFrames(model, torch.tensor(xs, dtype=dtype, device=device),
                           torch.tensor(cells, dtype=dtype, device=device))


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-9), (torch.float32, 1e-4)])
def test_gradient_matches_differences_and_autograd(device, dtype, tol):
    xs, cells, table, labels = _problem()
    # the oracle sees the positions the GPU sees
    xs = torch.tensor(xs, dtype=dtype).double().numpy()
    frames = _frames(xs, cells, table, dtype, device)
    scales = fit.scales(**{k: torch.tensor(v) for k, v in labels.items()})
    t = torch.tensor(table[None], dtype=dtype, device=device, requires_grad=True)
    pred = frames.predict(t)
    ref = _oracle_predict(xs, cells, table)
    fp_tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert _rel(pred["forces"].detach().cpu().numpy(), ref["forces"]) <= 10 * fp_tol
    assert _rel(pred["virial"].detach().cpu().numpy(), ref["virial"]) <= 10 * fp_tol
    lab = {k: torch.tensor(v, dtype=dtype, device=device) for k, v in labels.items()}
    value, _ = fit.loss(pred, lab, WEIGHTS, scales)
    value.backward()
    grad = t.grad[0].double().cpu().numpy()
    assert grad[-1] == 0.0

    auto = _autograd_gradient(xs, cells, table, labels, scales)
    if dtype == torch.float32:
        # no closer to fp64 than fp32 allows: the reference's own route in fp32 sits 1.8e-4
        # from fp64 on these frames (the knot coordinate (r - start) / h loses fp32 bits)
        route = _autograd_gradient(xs, cells, table, labels, scales, dtype=torch.float32)
        tol = max(tol, 1.5 * _rel(route, auto))
    assert _rel(grad, auto) <= tol

    # central differences of the oracle loss on knots spread over the table
    knots = np.unique(np.concatenate([np.arange(0, KNOTS, 11), [KNOTS - 2, KNOTS - 1]]))
    h = 0.25
    fd = np.zeros(len(knots))
    for m, k in enumerate(knots):
        up, down = table.copy(), table.copy()
        up[k] += h
        down[k] -= h
        if k == KNOTS - 1:  # the last knot is read as 0 whatever it holds
            up[k] = down[k] = 0.0
        fd[m] = (_np_loss(_oracle_predict(xs, cells, up), labels, scales) -
                 _np_loss(_oracle_predict(xs, cells, down), labels, scales)) / (2 * h)
    assert abs(fd[-1]) == 0.0
    scale = np.abs(auto).max()
    assert np.abs(fd - auto[knots]).max() <= 1e-8 * scale  # the two fp64 references agree
    assert np.abs(grad[knots] - fd).max() <= tol * scale


def test_repeated_backward_is_deterministic(device):
    xs, cells, table, labels = _problem(frames=2, beads=512)
    frames = _frames(xs, cells, table, torch.float32, device)
    scales = fit.scales(**{k: torch.tensor(v) for k, v in labels.items()})
    lab = {k: torch.tensor(v, dtype=torch.float32, device=device) for k, v in labels.items()}
    grads = []
    for _ in range(2):
        t = torch.tensor(table[None], dtype=torch.float32, device=device, requires_grad=True)
        value, _ = fit.loss(frames.predict(t), lab, WEIGHTS, scales)
        value.backward()
        grads.append(t.grad.clone())
    assert torch.isfinite(grads[0]).all()
    # fp64 atomics in LDS sum in any order: equal to rounding, not bit for bit
    assert torch.allclose(grads[0], grads[1], rtol=1e-6, atol=1e-9 * float(grads[0].abs().max()))


def test_full_workload_gradient(device):
    """The force-matching workload at its stated size: 64 frames of 4,096 beads (8 nm^-3), the
    200-knot table, fp32 on the GPU. dLoss/dknots against the fp64 autograd restatement and the
    fp32 route (the same bar as the small frames: max(1e-4, 1.5 x the fp32 route's distance from
    fp64)), and against fp64 central differences of the oracle's loss on three knots and one
    random direction (exact up to rounding: the loss is quadratic in the knots)."""
    xs, cells, table, labels = _problem(frames=64, beads=4096)
    xs = torch.tensor(xs, dtype=torch.float32).double().numpy()
    frames = _frames(xs, cells, table, torch.float32, device)
    scales = fit.scales(**{k: torch.tensor(v) for k, v in labels.items()})
    t = torch.tensor(table[None], dtype=torch.float32, device=device, requires_grad=True)
    lab = {k: torch.tensor(v, dtype=torch.float32, device=device) for k, v in labels.items()}
    value, _ = fit.loss(frames.predict(t), lab, WEIGHTS, scales)
    value.backward()
    grad = t.grad[0].double().cpu().numpy()
    assert grad[-1] == 0.0
    auto = _autograd_gradient(xs, cells, table, labels, scales)
    route = _autograd_gradient(xs, cells, table, labels, scales, dtype=torch.float32)
    tol = max(1e-4, 1.5 * _rel(route, auto))
    print(f"64 x 4096: gradient vs fp64 autograd {_rel(grad, auto):.2e}, vs the fp32 route "
          f"{_rel(grad, route):.2e}; the fp32 route vs fp64 {_rel(route, auto):.2e}", flush=True)
    assert _rel(grad, auto) <= tol

    pairs = [_pairs(on.wrap(x, c)[0], c) for x, c in zip(xs, cells)]
    spacing = np.array([(CUTOFF - START) / (KNOTS - 1)])

    def loss_of(tab):
        pred = {"energy": [], "forces": [], "virial": []}
        for x, c, p in zip(xs, cells, pairs):
            e, f, w = op.table_efw(on.wrap(x, c)[0], c, p, np.zeros(len(x), np.int64), tab[None],
                                   np.array([START]), CUTOFF, spacing=spacing)
            pred["energy"].append(e)
            pred["forces"].append(f)
            pred["virial"].append(w)
        return _np_loss({k: np.stack(v) for k, v in pred.items()}, labels, scales)

    rng = np.random.default_rng(9)
    direction = rng.normal(size=KNOTS)
    direction[-1] = 0.0
    h = 0.25
    scale = np.abs(auto).max()
    for k in (3, 97, 180):
        unit = np.zeros(KNOTS)
        unit[k] = 1.0
        fd = (loss_of(table + h * unit) - loss_of(table - h * unit)) / (2 * h)
        assert abs(fd - auto[k]) <= 1e-8 * scale, (k, fd, auto[k])
        assert abs(grad[k] - fd) <= tol * scale, (k, grad[k], fd)
    fd = (loss_of(table + h * direction) - loss_of(table - h * direction)) / (2 * h)
    assert abs(fd - auto @ direction) <= 1e-8 * scale * np.abs(direction).sum()
    assert abs(grad @ direction - fd) <= tol * scale * np.abs(direction).sum()
