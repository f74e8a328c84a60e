"""``nabladyn_amd.training`` on the GPU against the oracle's restatement of the reference's
double backward (``oracle/training.py``; ``training/_loss.py:71-125``, ``_fit.py:33-70``).

- The batch gradient of typed LJ (sigma and epsilon), typed tables, a sum of both, and each
  channel alone, from the parameter VJP kernel, equals the oracle's autograd gradient on the
  same (rounded) positions: fp64 within 1e-9 of the largest entry; fp32 within
  max(1e-5, 1.5 x the distance of the oracle's own route run in fp32 from fp64).
- The pair stacks hold each frame's pair set (brute force), self-pairs after it, at the width
  ``width`` gives, including the recount when a frame overflows the first count.
- One SGD update moves the trained rows by -lr x gradient and leaves held rows as they were.
- ``fit`` recovers a hidden LJ's sigma and epsilon from labels made with it.
- A model outside the fused kernels trains through autograd, as in the reference.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import potential, training
from nabladyn_amd.training import _fit
from oracle import md as omd
from oracle import neighbors as on
from oracle import potentials as op
from oracle import training as ot

pytestmark = pytest.mark.gpu

ATOMS, LENGTH = 216, 2.4
WEIGHTS = {"energy": 1.0, "forces": 1.0, "virial": 0.5}
SIGMA = [0.33, 0.35, 0.37]
EPSILON = [0.9, 1.05, 1.2]
KNOTS, START = 30, 0.3


def _problem(dtype, frames=3, seed=11, cutoff=1.0):
    """Wrapped positions [F, N, 3] and cells [F, 3] in ``dtype`` (wrapped with the reference's
    arithmetic in that dtype), types, random labels, and the oracle's frames in fp64."""
    rng = np.random.default_rng(seed)
    types = rng.integers(0, 2, ATOMS)
    xs, cells, oracle_frames = [], [], []
    for f in range(frames):
        length = LENGTH + 0.05 * f
        cell = torch.full((3,), length, dtype=dtype)
        x = torch.tensor(on.jittered_lattice(6, length, 0.03, seed=seed + f), dtype=dtype)
        x, _ = omd.wrap(x, cell)
        xs.append(x)
        cells.append(cell)
        x64, c64 = x.double().numpy(), cell.double().numpy()
        index = torch.tensor(on.brute_pairs(x64, c64, cutoff).T).long()
        oracle_frames.append((torch.tensor(x64), torch.tensor(c64), index))
    labels = {"energy": rng.normal(-250, 20, frames), "forces": rng.normal(0, 40, (frames, ATOMS, 3)),
              "virial": rng.normal(0, 30, (frames, 3, 3))}
    labels = {k: torch.tensor(v, dtype=dtype) for k, v in labels.items()}
    return torch.stack(xs), torch.stack(cells), types, labels, oracle_frames


def _table_values(seed=6):
    rng = np.random.default_rng(seed)
    r = START + (1.0 - START) / (KNOTS - 1) * np.arange(KNOTS)
    base = np.clip(4.0 * ((0.36 / r) ** 12 - (0.36 / r) ** 6), -40, 40)
    table = np.stack([base + rng.normal(0, 0.1, KNOTS) for _ in range(3)])
    table[:, -1] = 0.0
    return table


def _models(kind, types, dtype, device, trainable=None):
    t = torch.tensor(types, device=device)
    lj_cutoff = 0.8 if kind == "sum" else 1.0

    def lj():
        return potential.lennard_jones(
            atom_types=t, sigma=torch.tensor(SIGMA, dtype=dtype, device=device),
            epsilon=torch.tensor(EPSILON, dtype=dtype, device=device), cutoff=lj_cutoff,
            trainable=trainable if trainable is not None else ("sigma", "epsilon"))

    def table():
        return potential.tabulated(
            atom_types=t, table=torch.tensor(_table_values(), dtype=dtype, device=device),
            start=torch.full((3,), START, dtype=dtype, device=device), cutoff=1.0,
            trainable=trainable if trainable is not None else ("table",))

    if kind == "lj":
        return lj()
    if kind == "table":
        return table()
    return potential.sum(lj(), table())


def _oracle_energy(kind, types, dtype=torch.float64):
    """The oracle's energy and its parameters (in ``dtype``) for the same model."""
    tt = torch.tensor(types)
    start = torch.full((3,), START, dtype=dtype)
    spacing = (1.0 - start) / (KNOTS - 1)
    lj_cutoff = 0.8 if kind == "sum" else 1.0
    params = []
    if kind in ("lj", "sum"):
        params += [torch.tensor(SIGMA, dtype=dtype), torch.tensor(EPSILON, dtype=dtype)]
    if kind in ("table", "sum"):
        params += [torch.tensor(_table_values(), dtype=dtype)]

    def energy(p, x, strain, cell, index):
        total = 0.0
        k = 0
        if kind in ("lj", "sum"):
            total = total + op.t_lj_energy(x, strain, cell, index, tt, p[0], p[1], lj_cutoff)
            k = 2
        if kind in ("table", "sum"):
            total = total + op.t_table_energy(x, strain, cell, index, tt, p[k], start, spacing,
                                              1.0)
        return total

    return energy, params


def _frames_in(frames, dtype):
    return [(x.to(dtype), c.to(dtype), i) for x, c, i in frames]


def _gradient(model, xs, cells, labels, weights, device, chunk=2):
    data = training.Dataset(xs, cells, **labels)
    channels = tuple(n for n in training.CHANNELS if n in weights)
    prepared = _fit._prepare(model, data, channels)
    trained = {n: p for n, p in model.named_parameters() if p.requires_grad}
    scales = training.scales(**{k: labels[k] for k in channels})
    gradient, value, _ = _fit.batch_gradient(model, trained, prepared, torch.arange(len(xs)),
                                             chunk, weights, scales)
    return gradient, float(value), scales


def _flat(values):
    return torch.cat([v.detach().double().cpu().reshape(-1) for v in values])


def _rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("kind", ["lj", "table", "sum"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batch_gradient_matches_oracle(device, kind, dtype):
    xs, cells, types, labels, frames = _problem(dtype)
    model = _models(kind, types, dtype, device)
    gradient, value, scales = _gradient(model, xs, cells, labels, WEIGHTS, device)
    ours = _flat(gradient.values())
    energy, params = _oracle_energy(kind, types)
    labels64 = {k: v.double() for k, v in labels.items()}
    ref_value, ref = ot.loss_and_gradient(energy, params, frames, labels64, WEIGHTS, scales)
    ref = _flat(ref)
    if dtype == torch.float64:
        tol = 1e-9
        assert abs(value - float(ref_value)) <= 1e-10 * abs(float(ref_value))
    else:
        energy32, params32 = _oracle_energy(kind, types, torch.float32)
        _, route = ot.loss_and_gradient(energy32, params32, _frames_in(frames, torch.float32),
                                        labels, WEIGHTS, scales)
        route_distance = _rel(_flat(route), ref)
        tol = max(1e-5, 1.5 * route_distance)
        print(f"{kind} fp32: ours {_rel(ours, ref):.2e}, oracle's fp32 route {route_distance:.2e}")
        assert abs(value - float(ref_value)) <= 1e-4 * abs(float(ref_value))
    assert _rel(ours, ref) <= tol, (_rel(ours, ref), tol)


@pytest.mark.parametrize("weights", [{"energy": 1.0}, {"forces": 2.0}, {"virial": 1.0},
                                     {"forces": 1.0, "virial": 3.0}])
@pytest.mark.parametrize("kind", ["lj", "table"])
def test_each_channel_alone(device, kind, weights):
    xs, cells, types, labels, frames = _problem(torch.float64, frames=2, seed=21)
    model = _models(kind, types, torch.float64, device)
    gradient, _, scales = _gradient(model, xs, cells, labels, weights, device, chunk=1)
    energy, params = _oracle_energy(kind, types)
    _, ref = ot.loss_and_gradient(energy, params, frames, labels, weights, scales)
    assert _rel(_flat(gradient.values()), _flat(ref)) <= 1e-9


def test_pair_stack_rows(device):
    dtype = torch.float64
    # the third frame is compressed to 2.2 nm: over twice the pairs of the first, which the
    # first count was sized for, so the count runs again sized at it
    lengths = [3.0, 2.6, 2.2, 2.8]
    xs, cells = [], []
    for f, length in enumerate(lengths):
        x = torch.tensor(on.jittered_lattice(6, length, 0.03, seed=30 + f), dtype=dtype)
        xs.append(x + 0.5 * length)  # outside the cell: the stack wraps
        cells.append(torch.full((3,), length, dtype=dtype))
    stack = training.pair_stack(torch.stack(xs).to(device), torch.stack(cells).to(device), 1.0)
    stack = stack.cpu()
    totals = []
    for f, (x, cell) in enumerate(zip(xs, cells)):
        xw = on.wrap(x.numpy(), cell.numpy())[0]
        expected = on.brute_pairs(xw, cell.numpy(), 1.0)
        totals.append(len(expected))
        row = stack[f].numpy()
        listed = row[:, row[0] != row[1]]
        assert np.array_equal(on.pair_set(listed, listed.shape[1]),
                              on.pair_set(expected.T, len(expected))), f
        pads = np.flatnonzero(row[0] == row[1])
        assert np.array_equal(row[0, pads], pads % ATOMS) and np.array_equal(row[1, pads],
                                                                             pads % ATOMS)
    assert totals[2] > 2 * totals[0]
    assert stack.shape[2] == training.width(stack.shape[2]) >= max(totals)


def test_sgd_update_moves_trained_rows_only(device):
    dtype = torch.float64
    xs, cells, types, labels, frames = _problem(dtype, frames=2, seed=41)
    model = _models("table", types, dtype, device, trainable={"table": [0, 2]})
    before = model.table.detach().clone()
    lr = 1e-3
    report = training.fit(model, training.Dataset(xs, cells, **labels),
                          torch.optim.SGD(model.parameters(), lr=lr), updates=1, batch_size=2,
                          energy=1.0, forces=1.0, virial=0.5)
    energy, params = _oracle_energy("table", types)
    scales = training.scales(**labels)
    _, (grad,) = ot.loss_and_gradient(energy, params, frames, labels, WEIGHTS, scales)
    moved = (model.table.detach() - before).cpu()
    assert torch.equal(moved[1], torch.zeros(KNOTS, dtype=dtype))
    assert torch.all(moved[:, -1] == 0)
    expected = -lr * grad[[0, 2]]
    assert float((moved[[0, 2]] - expected).abs().max()) <= 1e-9 * float(expected.abs().max())
    assert report.loss.shape == (1,) and set(report.training) == {"energy", "forces", "virial"}
    assert report.validation is None


def test_fit_recovers_a_hidden_lj(device, capsys):
    dtype = torch.float64
    frames, hidden = 8, (0.36, 1.1)
    xs, cells, energies, forces, virials = [], [], [], [], []
    types = np.zeros(ATOMS, np.int64)
    for f in range(frames):
        length = 2.3 + 0.03 * f
        x = on.wrap(on.jittered_lattice(6, length, 0.04, seed=50 + f), np.full(3, length))[0]
        cell = np.full(3, length)
        pairs = on.brute_pairs(x, cell, 1.0).T
        e, force, w = op.lj_efw(x, cell, pairs, types, [hidden[0]], [hidden[1]], 1.0)
        xs.append(x), cells.append(cell), energies.append(e), forces.append(force)
        virials.append(w)
    data = training.Dataset(np.stack(xs), np.stack(cells), energy=np.array(energies),
                            forces=np.stack(forces), virial=np.stack(virials))
    model = potential.lennard_jones(
        atom_types=torch.zeros(ATOMS, dtype=torch.int64, device=device),
        sigma=torch.tensor([0.345], dtype=dtype, device=device),
        epsilon=torch.tensor([0.95], dtype=dtype, device=device), cutoff=1.0,
        trainable=("sigma", "epsilon"))
    optimizer = torch.optim.Adam(model.parameters(), lr=1e-2)
    report = training.fit(model, data[:6], optimizer, updates=300, batch_size=3, forces=1.0,
                          virial=0.1, validation=data[6:], echo_every=100)
    printed = capsys.readouterr().out.splitlines()
    assert printed[0] == "update loss forces virial" and len(printed) == 4
    assert printed[-1].startswith("300 ")
    # the oracle's eager fit of the same problem ends at loss 1.8e-7 from 0.28
    assert report.loss[-1] < 1e-4 * report.loss[0]
    for errors in (report.training, report.validation):
        assert errors["forces"][1] < 0.02 * errors["forces"][0]
    assert abs(float(model.sigma.detach()) - hidden[0]) < 1e-3
    assert abs(float(model.epsilon.detach()) - hidden[1]) < 1e-2


class _Soft(potential.Potential):
    """A pair repulsion k (rc - r)^2 written in torch: not a fused model, so training
    differentiates it twice by autograd."""

    def __init__(self, n, k, cutoff, device):
        super().__init__()
        self.k = torch.nn.Parameter(torch.tensor([k], dtype=torch.float64, device=device))
        self._n, self._cutoff = n, cutoff

    cutoff = property(lambda self: self._cutoff)
    num_atoms = property(lambda self: self._n)
    dtype = property(lambda self: self.k.dtype)
    device = property(lambda self: self.k.device)

    def _energy_wrapped(self, positions, cell, pairs, strain=None):
        first, second = pairs.index[0].long(), pairs.index[1].long()
        from nabladyn_amd import space
        d = space.strained(space._displacement_wrapped(positions[first], positions[second], cell),
                           strain)
        r2 = d.square().sum(-1)
        valid = (first != second) & (r2 < self._cutoff ** 2)
        r = torch.where(valid, r2, 1.0).sqrt()
        return torch.where(valid, self.k * (self._cutoff - r).square(), 0.0).sum()


def test_generic_model_trains_through_autograd(device):
    xs, cells, types, labels, frames = _problem(torch.float64, frames=2, seed=61, cutoff=0.5)
    model = _Soft(ATOMS, 30.0, 0.5, device)
    gradient, _, scales = _gradient(model, xs, cells, labels, WEIGHTS, device)

    def energy(p, x, strain, cell, index):
        d = op._t_strained(op._t_displacement(x[index[0]], x[index[1]], cell), strain)
        r2 = d.square().sum(-1)
        valid = (index[0] != index[1]) & (r2 < 0.25)
        r = torch.where(valid, r2, 1.0).sqrt()
        return torch.where(valid, p[0] * (0.5 - r).square(), 0.0).sum()

    _, ref = ot.loss_and_gradient(energy, [torch.tensor([30.0], dtype=torch.float64)], frames,
                                  labels, WEIGHTS, scales)
    assert _rel(_flat(gradient.values()), _flat(ref)) <= 1e-9


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_frames_channels_match_each_frame_alone(device, dtype, tol):
    """E, F and W of a batch of frames from one launch, over the stacks (``ndn_pair_efw_frames``)
    and over the full rows built from them (``ndn_pair_efw_full_frames``), against each frame
    alone through ``quantity.energy_force_virial`` on its own list."""
    from nabladyn_amd import quantity
    from nabladyn_amd.neighbors import NeighborList

    xs, cells, types, _, _ = _problem(dtype, frames=3, seed=71)
    model = _models("sum", types, dtype, device)
    x, c = xs.to(device), cells.to(device)
    stacks = tuple(training.pair_stack(x, c, cutoff) for cutoff in model._list_cutoffs)
    rows = tuple(training._loss.frame_rows(stack, ATOMS) for stack in stacks)
    with torch.no_grad():
        over_stacks = training.predict(model, {}, x, c, stacks)
        over_rows = training._loss._predict(model, {}, x, c, stacks, training.CHANNELS,
                                            wrapped=False, rows=rows)
    efw = quantity.energy_force_virial(model)
    for f in range(3):
        nl = NeighborList(1.0, inner_cutoffs=(0.8,))
        nl.allocate(x[f], c[f])
        e, force, w = efw(x[f], c[f], nl.tensors())
        for batched in (over_stacks, over_rows):
            for got, want in ((batched["energy"][f], e), (batched["forces"][f], force),
                              (batched["virial"][f], w)):
                assert _rel(got.double().cpu(), want.double().cpu()) <= tol, f
