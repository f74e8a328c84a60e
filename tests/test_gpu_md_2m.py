"""The benched workload at its stated size: liquid-argon LJ NVT, 2,097,152 atoms (128^3 jittered
lattice, 46.3949 nm box), fp32, cutoff 1.0 nm, skin 0.2 nm, the default pruned list.

Checked through properties whose cost grows with the system, not with its square:
- the MD list against the oracle's C cell list at the list's own build frame, at the first
  build and after device-decided rebuilds: the total number of pairs and, for every atom, its
  neighbor count, the sum of its neighbors' ids and the sum of their squares (exact integers);
- 1,000 sampled rows against an O(N) brute-force search per row (an x-band prefilter, a strict
  superset, then the exact fp32 minimum-image test of ``traversal.cuh:31-38``);
- every step's rebuild and pruned-list refresh decision over 120 liquid steps against the
  oracle's reuse rule replayed on the frames the steps saw (bit-exact);
- the forces on the sampled atoms against an fp64 evaluation of the same fp32 positions:
  max|dF| / max|F| <= 1e-5.
"""

import ctypes

import numpy as np
import pytest
import torch

from nabladyn_amd import _native, md, potential, workloads
from oracle import neighbors as on
from test_gpu_md_decisions import record

pytestmark = pytest.mark.gpu

CUTOFF, SKIN, RADIUS = 1.0, 0.2, 1.2
SIGMA, EPSILON = workloads.ARGON_SIGMA, workloads.ARGON_EPSILON


@pytest.fixture(scope="module")
def simulation():
    device = torch.device("cuda", 0)
    x, cell = workloads.argon_lattice(128, seed=0)
    n = x.shape[0]
    masses = torch.full((n,), workloads.ARGON_MASS)
    v = md.maxwell_boltzmann(masses, 120.0, 1)
    model = potential.lennard_jones(
        atom_types=torch.zeros(n, dtype=torch.int64, device=device),
        sigma=torch.tensor([SIGMA], device=device), epsilon=torch.tensor([EPSILON], device=device),
        cutoff=CUTOFF)
    sim = md.nvt(model, torch.tensor(x, dtype=torch.float32, device=device), v.to(device),
                 masses.to(device), torch.tensor(cell, dtype=torch.float32, device=device),
                 timestep=0.004, temperature=120.0, thermostat_timescale=0.1, skin=SKIN)
    print(f"2M system: {n} atoms, bricks {sim._neighbors.brick}, stride {sim._neighbors.stride}, "
          f"halo capacity {sim._neighbors.halo_capacity}, pruned skin {sim._neighbors.inner_skin}",
          flush=True)
    return sim


def _rows(sim, rows):
    """The atom ids each of `rows` (slots) lists, as a [len(rows), stride] int64 tensor with -1
    past its neighbors."""
    nb = sim._neighbors
    n, stride = sim._num_atoms, nb.stride
    slots = torch.full((n * stride,), -1, dtype=torch.int32, device=sim._device)
    _native.call("md_list_entries", ctypes.byref(nb.native), ctypes.c_int64(n), 0,
                 _native.ptr(slots), _native.stream(sim._device))
    index = slots.view(n, stride)[rows].long()
    ids = nb.buffers["ids"].long()
    return torch.where(index >= 0, ids[index.clamp(min=0)], torch.full_like(index, -1))


def _signature(sim):
    """Per atom (in atom order): neighbor count, sum of neighbor ids, sum of their squares."""
    nb = sim._neighbors
    n, stride = sim._num_atoms, nb.stride
    slots = torch.full((n * stride,), -1, dtype=torch.int32, device=sim._device)
    _native.call("md_list_entries", ctypes.byref(nb.native), ctypes.c_int64(n), 0,
                 _native.ptr(slots), _native.stream(sim._device))
    index = slots.view(n, stride)
    ids = nb.buffers["ids"].long()
    out = torch.zeros((3, n), dtype=torch.int64, device=sim._device)
    for begin in range(0, n, 1 << 18):
        block = index[begin:begin + (1 << 18)].long()
        valid = block >= 0
        nbr = torch.where(valid, ids[block.clamp(min=0)], torch.zeros_like(block))
        atoms = ids[begin:begin + block.shape[0]]
        out[0, atoms] = valid.sum(dim=1)
        out[1, atoms] = nbr.sum(dim=1)
        out[2, atoms] = (nbr * nbr).sum(dim=1)
    del slots, index
    return out.cpu().numpy()


def _oracle_signature(x0, cell):
    b = on.build(x0, cell, RADIUS)
    assert b.status == 0
    i = b.pairs[0, :b.pair_total].astype(np.int64)
    j = b.pairs[1, :b.pair_total].astype(np.int64)
    n = x0.shape[0]
    count = np.bincount(i, minlength=n) + np.bincount(j, minlength=n)
    # integer sums below 2^53: exact in float64 whatever the order
    s1 = np.bincount(i, weights=j.astype(np.float64), minlength=n) + \
        np.bincount(j, weights=i.astype(np.float64), minlength=n)
    s2 = np.bincount(i, weights=(j * j).astype(np.float64), minlength=n) + \
        np.bincount(j, weights=(i * i).astype(np.float64), minlength=n)
    return b.pair_total, np.stack([count, s1.astype(np.int64), s2.astype(np.int64)])


def _check_list(sim, label):
    nb = sim._neighbors
    cell = sim.cell.cpu().numpy()
    x0 = sim._to_atoms(nb.reference_positions).cpu().numpy()
    total, want = _oracle_signature(x0, cell)
    got = _signature(sim)
    print(f"{label}: {total} pairs within {RADIUS} nm, rebuild {nb._rebuild_count}", flush=True)
    assert int(got[0].sum()) == 2 * total
    np.testing.assert_array_equal(got, want)


def _near(x, centers, cell, radius):
    """For each center, the atoms within `radius` (exact fp32 minimum-image test, the list's
    expression) and their fp32 displacement vectors, found from an x-band superset."""
    length = cell[0]
    order = np.argsort(x[:, 0], kind="stable")
    xs = x[order, 0]
    margin = np.float32(radius) + np.float32(0.01)
    r2 = np.float32(radius) * np.float32(radius)
    out = []
    for c in centers:
        lo, hi = x[c, 0] - margin, x[c, 0] + margin
        spans = [(lo, hi)]
        if lo < 0:
            spans.append((lo + length, length))
        if hi >= length:
            spans.append((0, hi - length))
        cand = np.unique(np.concatenate(
            [order[np.searchsorted(xs, a, "left"):np.searchsorted(xs, b, "right")] for a, b in spans]))
        d = np.stack([on.minimum_image_np(np.full(len(cand), x[c, k]), x[cand, k], cell[k])
                      for k in range(3)], axis=1)
        sq = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        hit = (sq < r2) & (cand != c)
        out.append((cand[hit], d[hit]))
    return out


def test_2m_list_decisions_rows_and_forces(simulation):
    sim = simulation
    nb = sim._neighbors
    _check_list(sim, "first build")
    sim.run(100)  # the lattice melts; rebuilds start
    rows = record(sim, 120)
    print(f"120 liquid steps: {rows[:, 0].sum()} rebuilds, {rows[:, 1].sum()} refreshes", flush=True)
    assert rows[:, 0].sum() >= 2
    np.testing.assert_array_equal(rows[:, 0], rows[:, 2])
    np.testing.assert_array_equal(rows[:, 1], rows[:, 3])
    _check_list(sim, "after device-decided rebuilds")

    # sampled rows against brute force at the list's build frame
    cell = sim.cell.cpu().numpy()
    rng = np.random.default_rng(0)
    centers = rng.choice(sim._num_atoms, 1000, replace=False)
    x0 = sim._to_atoms(nb.reference_positions).cpu().numpy()
    slot_of = torch.empty_like(nb.buffers["ids"])
    slot_of[nb.buffers["ids"].long()] = torch.arange(sim._num_atoms, dtype=torch.int32,
                                                     device=sim._device)
    listed = _rows(sim, slot_of[torch.as_tensor(centers, device=sim._device)].long()).cpu().numpy()
    for k, (hits, _) in enumerate(_near(x0, centers, cell, RADIUS)):
        mine = np.sort(listed[k][listed[k] >= 0])
        np.testing.assert_array_equal(mine, np.sort(hits))

    # forces on the sampled atoms at the current frame, fp64 on the same fp32 positions
    x = sim._to_atoms(sim.positions).cpu().numpy()
    forces = sim._to_atoms(sim.forces).cpu().numpy()
    want = np.zeros((len(centers), 3))
    # the model's fp32 parameters, in fp64
    s2 = np.float64(np.float32(SIGMA)) ** 2
    eps = np.float64(np.float32(EPSILON))
    for k, (hits, d) in enumerate(_near(x, centers, cell, CUTOFF)):
        d64 = d.astype(np.float64)
        r2 = (d64 * d64).sum(axis=1)
        p6 = (s2 / r2) ** 3
        want[k] = ((24.0 * eps * (2.0 * p6 * p6 - p6) / r2)[:, None] * d64).sum(axis=0)
    err = np.abs(forces[centers] - want).max() / np.abs(want).max()
    print(f"sampled forces: max|dF|/max|F| = {err:.2e} against fp64", flush=True)
    assert err <= 1e-5
