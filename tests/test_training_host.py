"""The host side of ``nabladyn_amd.training`` on the CPU: the Dataset's checks and indexing
(``training/_dataset.py``), the stack width, the batch order, held entries in the optimizer
step, the channel scales and the loss (``training/_loss.py:26-44,105-125``), and the argument
checks of ``fit`` (``training/_fit.py:103-171``), which all run before anything touches a GPU.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import potential, training
from nabladyn_amd.training import _fit


def _data(frames=5, atoms=8, dtype=np.float64, seed=0):
    rng = np.random.default_rng(seed)
    return dict(positions=rng.uniform(0, 3, (frames, atoms, 3)).astype(dtype),
                cells=np.full((frames, 3), 3.0, dtype),
                energy=rng.normal(size=frames).astype(dtype),
                forces=rng.normal(size=(frames, atoms, 3)).astype(dtype),
                virial=rng.normal(size=(frames, 3, 3)).astype(dtype))


def test_dataset_copies_and_indexes():
    fields = _data()
    data = training.Dataset(**fields)
    fields["positions"][0, 0, 0] = 99.0
    assert float(data.positions[0, 0, 0]) != 99.0
    assert len(data) == 5 and list(data.labels) == ["energy", "forces", "virial"]
    assert torch.equal(data[1].positions, data.positions[1:2])
    assert torch.equal(data[1:4].labels["energy"], data.labels["energy"][1:4])
    picked = data[np.array([4, 0])]
    assert torch.equal(picked.cells, data.cells[[4, 0]])
    assert list(training.Dataset(fields["positions"], fields["cells"]).labels) == []


@pytest.mark.parametrize("change, message", [
    (dict(positions=np.zeros((0, 8, 3))), "at least one frame"),
    (dict(positions=np.zeros((5, 8, 2))), "at least one frame"),
    (dict(cells=np.ones((5, 2))), r"cells must be \[5, 3\]"),
    (dict(forces=np.zeros((5, 7, 3))), r"forces must be \[5, 8, 3\]"),
    (dict(energy=np.zeros(5, np.float32)), "one dtype"),
    (dict(virial=np.full((5, 3, 3), np.nan)), "virial must be finite"),
    (dict(cells=np.zeros((5, 3))), "cells must be positive"),
])
def test_dataset_rejects(change, message):
    fields = dict(_data(), **change)
    with pytest.raises(ValueError, match=message):
        training.Dataset(**fields)


def test_dataset_rejects_integer_fields():
    fields = _data()
    fields = {k: (v * 0).astype(np.int64) + 1 for k, v in fields.items()}
    with pytest.raises(ValueError, match="one dtype"):
        training.Dataset(**fields)


def test_width_rounds_to_an_eighth_of_the_leading_power():
    assert [training.width(c) for c in (1, 7, 15, 16, 17, 1000, 4096, 4097)] == \
        [1, 7, 15, 16, 18, 1024, 4096, 4608]
    for count in range(1, 5000, 37):
        w = training.width(count)
        assert count <= w <= count * 1.125 + 1


def test_batches_shuffle_each_epoch():
    order = training.batches(10, 4, 7, seed=3)
    assert [len(b) for b in order] == [4, 4, 2, 4, 4, 2, 4]
    assert sorted(torch.cat(order[:3]).tolist()) == list(range(10))
    assert sorted(torch.cat(order[3:6]).tolist()) == list(range(10))
    assert all(torch.equal(a, b) for a, b in zip(order, training.batches(10, 4, 7, seed=3)))
    assert not all(torch.equal(a, b) for a, b in zip(order, training.batches(10, 4, 7, seed=4)))


def test_optimizer_step_keeps_held_rows():
    model = potential.lennard_jones(
        atom_types=torch.zeros(4, dtype=torch.int64), sigma=torch.tensor([0.3, 0.4, 0.5]),
        epsilon=torch.tensor([1.0, 2.0, 3.0]), cutoff=1.0, trainable={"sigma": [0, 2]})
    trained = {n: p for n, p in model.named_parameters() if p.requires_grad}
    assert list(trained) == ["sigma"]
    held = _fit.held_entries(model, trained)
    optimizer = torch.optim.AdamW(trained.values(), lr=0.1, weight_decay=0.5)
    _fit.optimizer_step(optimizer, trained, {"sigma": torch.ones(3)}, held)
    sigma = model.sigma.detach()
    assert float(sigma[1]) == float(torch.tensor(0.4))  # bit for bit
    assert float(sigma[0]) != pytest.approx(0.3) and float(sigma[2]) != pytest.approx(0.5)


def test_scales_and_loss():
    energy = torch.tensor([1.0, 3.0])
    forces = torch.tensor([[[3.0, 4.0, 0.0]], [[0.0, 0.0, 0.0]]])
    virial = torch.full((2, 3, 3), 2.0)
    s = training.scales(energy=energy, forces=forces, virial=virial)
    assert s["energy"] == pytest.approx(1.0)                  # about the mean
    assert s["forces"] == pytest.approx(np.sqrt(25.0 / 6))    # plain RMS
    assert s["virial"] == pytest.approx(2.0)                  # constant: the plain RMS
    assert training.scales(energy=torch.zeros(3))["energy"] == 1.0
    predicted = {"energy": torch.tensor([2.0, 3.0]), "forces": forces + 1.0}
    labels = {"energy": energy, "forces": forces}
    weights = {"energy": 2.0, "forces": 0.5}
    value, terms = training.loss(predicted, labels, weights, {"energy": 1.0, "forces": 2.0})
    assert float(terms["energy"]) == pytest.approx(0.5)       # (1^2 + 0) / 2
    assert float(terms["forces"]) == pytest.approx(0.25)      # every component off by 1/2
    assert float(value) == pytest.approx(2.0 * 0.5 + 0.5 * 0.25)
    weighted, terms = training.loss(predicted, labels, weights, {"energy": 1.0, "forces": 2.0},
                                    torch.tensor([1.0, 0.0]))
    assert float(terms["energy"]) == pytest.approx(1.0)


def _lj(trainable=("sigma",), atoms=8):
    return potential.lennard_jones(atom_types=torch.zeros(atoms, dtype=torch.int64),
                                   sigma=torch.tensor([0.3], dtype=torch.float64),
                                   epsilon=torch.tensor([1.0], dtype=torch.float64), cutoff=1.0,
                                   trainable=trainable)


def test_fit_checks_its_arguments():
    data = training.Dataset(**_data())
    sgd = torch.optim.SGD([torch.zeros(1, requires_grad=True)], lr=0.1)
    with pytest.raises(TypeError, match="Potential"):
        training.fit(torch.nn.Linear(1, 1), data, sgd, updates=1, batch_size=1, forces=1.0)
    with pytest.raises(ValueError, match="at least one channel"):
        training.fit(_lj(), data, sgd, updates=1, batch_size=1)
    with pytest.raises(ValueError, match="finite positive"):
        training.fit(_lj(), data, sgd, updates=1, batch_size=1, forces=0.0)
    with pytest.raises(ValueError, match="batch_size between"):
        training.fit(_lj(), data, sgd, updates=1, batch_size=6, forces=1.0)
    with pytest.raises(ValueError, match="nonnegative integer"):
        training.fit(_lj(), data, sgd, updates=1.5, batch_size=1, forces=1.0)
    with pytest.raises(ValueError, match="no trainable parameter"):
        training.fit(_lj(trainable=()), data, sgd, updates=1, batch_size=1, forces=1.0)
    with pytest.raises(ValueError, match="hold 8 atoms, the model 9"):
        training.fit(_lj(atoms=9), data, sgd, updates=1, batch_size=1, forces=1.0)
    unlabeled = training.Dataset(**{k: v for k, v in _data().items() if k != "virial"})
    with pytest.raises(ValueError, match="no virial labels"):
        training.fit(_lj(), unlabeled, sgd, updates=1, batch_size=1, virial=1.0)


def test_frame_rows_list_every_pair_both_ways_in_stack_order():
    """The full rows built from a stack [F, 2, C]: row r of frame f holds, in the stack's column
    order, the second atom of each pair with r first, then the first atom of each pair with r
    second; self-pairs (the padding) are left out; slices are 64 rows wide, column-major."""
    from nabladyn_amd.training._loss import frame_rows

    rng = np.random.default_rng(3)
    frames, atoms, columns = 3, 150, 700
    stack = torch.tensor(rng.integers(0, atoms, (frames, 2, columns)), dtype=torch.int32)
    stack[:, :, -40:] = torch.arange(40, dtype=torch.int32) % atoms  # padding
    rows = frame_rows(stack, atoms)
    per_frame = -(-atoms // 64)
    assert rows.counts.shape == (frames, atoms) and rows.slices.shape == (frames, per_frame + 1)
    for f in range(frames):
        first, second = stack[f, 0].tolist(), stack[f, 1].tolist()
        for r in range(atoms):
            want = [b for a, b in zip(first, second) if a == r and a != b]
            want += [a for a, b in zip(first, second) if b == r and a != b]
            count = int(rows.counts[f, r])
            base = int(rows.slices[f, r // 64]) + r % 64
            got = [int(rows.index[f, base + 64 * k]) for k in range(count)]
            assert got == want, (f, r)
        widths = rows.counts[f].tolist() + [0] * (per_frame * 64 - atoms)
        for s in range(per_frame):
            assert int(rows.slices[f, s + 1] - rows.slices[f, s]) == 64 * max(
                widths[64 * s:64 * s + 64])
