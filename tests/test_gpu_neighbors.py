"""Parity of the HIP neighbor list with the oracle: the pair index [2, C] bit for bit (column
order and padding included), pair totals, capacities, rebuild decisions over position
sequences, the full list, and the latched status bits."""

import math

import numpy as np
import pytest
import torch

from nabladyn_amd import space
from nabladyn_amd.neighbors import NeighborList, ops
from oracle import neighbors as on

pytestmark = pytest.mark.gpu

DTYPES = [(torch.float32, np.float32), (torch.float64, np.float64)]


def _allocate(x, cell, cutoff, skin, dtype, device, **kw):
    nl = NeighborList(cutoff, skin=skin, **kw)
    nl.allocate(torch.tensor(x, dtype=dtype, device=device), torch.tensor(cell, dtype=dtype, device=device))
    return nl


def _expected(x, cell, radius, npdt, multiplier=1.25):
    xw, _ = on.wrap(x.astype(npdt), cell.astype(npdt))
    return on.build(xw, cell.astype(npdt), radius, capacity_multiplier=multiplier)


@pytest.mark.parametrize("dtype,npdt", DTYPES)
def test_argon_512_documented_counts_and_exact_index(argon, device, dtype, npdt):
    nl = _allocate(argon["positions"], argon["cell"], 1.0, 0.1, dtype, device)
    assert nl.num_pairs == 29014
    assert nl.pair_capacity == 36268
    assert nl.all_pairs
    ref = _expected(argon["positions"], argon["cell"], 1.1, npdt)
    got = nl.tensors().index.cpu().numpy()
    np.testing.assert_array_equal(got, ref.pairs)


def test_wrap_matches_oracle_bitwise(device):
    rng = np.random.default_rng(5)
    cell = np.array([3.1, 4.7, 5.3])
    x = rng.uniform(-20, 20, (5000, 3))
    x[:10] = cell * rng.integers(-3, 3, (10, 3))  # exact multiples
    x[10:20] = -1e-9
    for dtype, npdt in DTYPES:
        got, crossings = space.wrap(torch.tensor(x, dtype=dtype, device=device),
                                    torch.tensor(cell, dtype=dtype, device=device))
        ref, ref_cross = on.wrap(x.astype(npdt), cell.astype(npdt))
        np.testing.assert_array_equal(got.cpu().numpy(), ref)
        np.testing.assert_array_equal(crossings.cpu().numpy(), ref_cross)


GRID_CASES = [
    # (N, cell, cutoff, skin): 27-bin path, mixed (short axis) path, anisotropic
    (6000, (6.1, 5.3, 7.7), 1.0, 0.2),
    (4000, (2.5, 6.0, 6.0), 1.0, 0.2),
    (3000, (9.0, 4.0, 3.7), 0.9, 0.15),
    (20000, (10.0, 10.0, 10.0), 1.0, 0.2),
]


@pytest.mark.parametrize("case", GRID_CASES)
@pytest.mark.parametrize("dtype,npdt", DTYPES)
def test_grid_path_exact_index(device, case, dtype, npdt):
    n, cell, cutoff, skin = case
    cell = np.array(cell)
    rng = np.random.default_rng(n)
    x = rng.uniform(0, 1, (n, 3)) * cell + rng.normal(0, 3, (n, 3))  # many outside the cell
    nl = _allocate(x, cell, cutoff, skin, dtype, device)
    ref = _expected(x, cell, cutoff + skin, npdt)
    assert not nl.all_pairs
    assert nl.num_pairs == ref.pair_total
    got = nl.tensors().index.cpu().numpy()
    np.testing.assert_array_equal(got, ref.pairs)
    brute = on.brute_pairs(on.wrap(x.astype(npdt), cell.astype(npdt))[0], cell.astype(npdt), cutoff + skin)
    np.testing.assert_array_equal(on.pair_set(got, nl.num_pairs), brute)


@pytest.mark.parametrize("dtype,npdt", DTYPES)
def test_lattice_config1(device, dtype, npdt):
    x = on.jittered_lattice(8, 2.90, 0.02, seed=0)
    cell = np.full(3, 2.90)
    nl = _allocate(x, cell, 1.0, 0.1, dtype, device)
    ref = _expected(x, cell, 1.1, npdt)
    np.testing.assert_array_equal(nl.tensors().index.cpu().numpy(), ref.pairs)


def _displacements(rng, n, scale):
    return rng.normal(0, scale, (n, 3))


@pytest.mark.parametrize("dtype,npdt", DTYPES)
@pytest.mark.parametrize("n,cell", [(512, 2.9), (8000, 5.8)])
def test_update_sequence_rebuild_decisions(device, dtype, npdt, n, cell):
    cutoff, skin = 1.0, 0.1
    per = round(n ** (1 / 3))
    x = on.jittered_lattice(per, cell, 0.02, seed=1)
    L = np.full(3, cell)
    nl = _allocate(x, L, cutoff, skin, dtype, device)
    from oracle.md import CPUList

    ref = CPUList(cutoff, skin)
    ref.allocate(on.wrap(x.astype(npdt), L.astype(npdt))[0], L.astype(npdt))
    rng = np.random.default_rng(2)
    cur = x.copy()
    decisions = []
    for step in range(25):
        cur = cur + _displacements(rng, n, 0.004 * (1 + step % 7))
        xd = torch.tensor(cur, dtype=dtype, device=device)
        Ld = torch.tensor(L, dtype=dtype, device=device)
        before = nl._rebuild_count
        nl.update(xd, Ld)
        rebuilt = nl._rebuild_count != before
        xw = on.wrap(cur.astype(npdt), L.astype(npdt))[0]
        expected = ref.update(xw, L.astype(npdt))
        decisions.append((rebuilt, expected))
        assert rebuilt == expected, f"step {step}"
        np.testing.assert_array_equal(nl.tensors().index.cpu().numpy(), ref.index)
    assert any(d[0] for d in decisions) and not all(d[0] for d in decisions)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_full_list_is_symmetric_closure(device, dtype):
    rng = np.random.default_rng(11)
    for n, cell in ((700, (3.0, 3.0, 3.0)), (9000, (7.0, 6.5, 8.0))):
        cell = np.array(cell)
        x = rng.uniform(0, 1, (n, 3)) * cell
        nl = _allocate(x, cell, 1.0, 0.2, dtype, device, full=True)
        half = on.pair_set(nl.tensors().index.cpu().numpy(), nl.num_pairs)
        full = nl._device.full
        counts = full["counts"].cpu().numpy()
        slices = full["slices"].cpu().numpy()
        index = full["index"].cpu().numpy()
        rows = (nl._device.buffers["sorted_ids"].cpu().numpy() if not nl.all_pairs else np.arange(n))
        got = []
        for r in range(n):
            base = slices[r // 64] + r % 64
            nbrs = index[base + 64 * np.arange(counts[r])]
            got.append(np.stack([np.full(counts[r], rows[r]), nbrs], axis=1))
        got = np.concatenate(got)
        sym = np.concatenate([half, half[:, ::-1]])
        key = lambda a: a[np.lexsort((a[:, 1], a[:, 0]))]
        np.testing.assert_array_equal(key(got), key(sym))


def test_forced_rebuild_and_torch_ops_path(device):
    x = on.jittered_lattice(20, 7.0, 0.02, seed=3)
    L = np.full(3, 7.0)
    nl = _allocate(x, L, 1.0, 0.2, torch.float32, device)
    xd, _ = space.wrap(torch.tensor(x, dtype=torch.float32, device=device),
                       torch.tensor(L, dtype=torch.float32, device=device))
    Ld = torch.tensor(L, dtype=torch.float32, device=device)
    nl._rebuild_wrapped(xd, Ld)
    nl._check_status()
    assert nl._rebuild_count == 2
    # the same update through torch.ops.nabladyn_neighbors.update
    dev = nl._device
    r = dev.reuse
    r["reference_cell"].fill_(math.inf)
    ops().update(xd, Ld, 1.0, 1.2, r["reference_positions"], r["reference_cell"], r["limits"],
                 r["parameters"], r["rebuild"], r["rebuild_count"], *dev.args())
    torch.cuda.synchronize()
    assert nl._rebuild_count == 3
    ref = _expected(x, L, 1.2, np.float32)
    np.testing.assert_array_equal(nl.tensors().index.cpu().numpy(), ref.pairs)


def test_status_bits(device):
    x = on.jittered_lattice(10, 4.0, 0.02, seed=4)
    L = torch.full((3,), 4.0, device=device)
    xd = torch.tensor(x, dtype=torch.float32, device=device)
    nl = NeighborList(1.0, skin=0.1)
    with pytest.raises(RuntimeError, match="allocate"):
        nl.update(xd, L)
    nl.allocate(xd, L)
    bad = xd.clone()
    bad[7, 1] = float("nan")
    with pytest.raises(ValueError, match="finite"):
        nl.update(bad, L)
    with pytest.raises(RuntimeError, match="failed"):
        nl.tensors()
    nl.allocate(xd, L)
    with pytest.raises(ValueError, match="twice the cutoff"):
        nl.update(xd, torch.full((3,), 1.9, device=device))
    nl2 = NeighborList(1.0, skin=0.1, capacity_multiplier=1.0)
    nl2.allocate(xd, L)
    squeezed, small = xd * 0.9, L * 0.9
    with pytest.raises(ValueError, match="pair capacity") as err:
        nl2.update(squeezed, small)
    small_np = small.cpu().numpy()
    needed = on.build(on.wrap(squeezed.cpu().numpy(), small_np)[0], small_np, 1.1).pair_total
    assert str(needed) in str(err.value)
    with pytest.raises(ValueError, match="twice the cutoff"):
        NeighborList(1.0).allocate(xd, torch.full((3,), 2.0, device=device))


def test_repeated_builds_are_identical(device):
    rng = np.random.default_rng(9)
    cell = np.array([8.0, 8.0, 8.0])
    x = rng.uniform(0, 8, (10000, 3))
    a = _allocate(x, cell, 1.0, 0.2, torch.float32, device)
    b = _allocate(x, cell, 1.0, 0.2, torch.float32, device)
    assert torch.equal(a.tensors().index, b.tensors().index)
