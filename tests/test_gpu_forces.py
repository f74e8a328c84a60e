"""Parity of the HIP energy/force/virial kernels with the fp64 oracle, the reference's fp64 LJ
labels (lj_frames.npz) and the documented argon_512 numbers.

Tolerance (stated for fp32, the reference's MD dtype): LJ max|dF| / max|F_ref| <= 1e-5,
|dW| / max|W_ref| <= 1e-5 and |dE| / max(1, |E|) <= 1e-5 against an fp64 evaluation of the
same fp32 positions on the same pairs. Tables in fp32: the same 1e-5 against the reference's
own fp32 route (the oracle's restatement of its autograd evaluation), with each result's fp64
error printed beside it. fp64 results: 1e-12 (tables 3e-12).
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import potential, quantity
from nabladyn_amd.neighbors import NeighborList
from oracle import neighbors as on
from oracle import potentials as op
from _parity import check_against_route, table_route

pytestmark = pytest.mark.gpu

TOL = {torch.float32: 1e-5, torch.float64: 1e-12}


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_lj_frames_labels(lj_frames, device, dtype):
    d = lj_frames
    types = torch.tensor(d["atom_types"], device=device)
    model = potential.lennard_jones(
        atom_types=types, sigma=torch.tensor(d["sigma"], dtype=dtype, device=device),
        epsilon=torch.tensor(d["epsilon"], dtype=dtype, device=device), cutoff=1.0)
    efw = quantity.energy_force_virial(model)
    for f in range(d["positions"].shape[0]):
        x = torch.tensor(d["positions"][f], dtype=dtype, device=device)
        L = torch.tensor(d["cells"][f], dtype=dtype, device=device)
        nl = NeighborList(1.0, skin=0.1)
        nl.allocate(x, L)
        e, F, W = efw(x, L, nl.tensors())
        tol = TOL[dtype] if dtype == torch.float64 else 5e-5  # fp32 positions vs fp64 labels
        if dtype == torch.float64:
            assert abs(float(e) - d["energy"][f]) <= 1e-10 * abs(d["energy"][f])
            assert _rel(F.cpu().numpy(), d["forces"][f]) <= tol
            assert _rel(W.cpu().numpy(), d["virial"][f]) <= tol
        # against fp64 on the same (possibly fp32-rounded) positions and pairs
        xw = on.wrap(x.cpu().numpy(), L.cpu().numpy())[0]
        E0, F0, W0 = op.lj_efw(xw, L.cpu().numpy(), nl.tensors().index.cpu().numpy(),
                               d["atom_types"], d["sigma"].astype(x.cpu().numpy().dtype),
                               d["epsilon"].astype(x.cpu().numpy().dtype), 1.0)
        assert _rel(F.cpu().numpy(), F0) <= TOL[dtype]
        assert _rel(W.cpu().numpy(), W0) <= TOL[dtype]
        assert abs(float(e) - E0) <= TOL[dtype] * max(1.0, abs(E0))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_argon_documented_numbers(argon, device, dtype):
    x = torch.tensor(argon["positions"], dtype=dtype, device=device)
    L = torch.tensor(argon["cell"], dtype=dtype, device=device)
    v = torch.tensor(argon["velocities"], dtype=dtype, device=device)
    model = potential.lennard_jones(
        atom_types=torch.zeros(512, dtype=torch.int64, device=device),
        sigma=torch.tensor([0.34], dtype=dtype, device=device),
        epsilon=torch.tensor([0.996], dtype=dtype, device=device), cutoff=1.0)
    nl = NeighborList(1.0, skin=0.1)
    nl.allocate(x, L)
    e, F, W = quantity.energy_force_virial(model)(x, L, nl.tensors())
    from nabladyn_amd import observables, units
    K = observables.kinetic_energy(v, torch.full((512,), 39.948, dtype=dtype, device=device))
    P = float(observables.pressure(K, W, L)) / units.BAR
    assert f"{float(e):.2f}" == "-2526.17"
    assert f"{float(F.norm(dim=1).max()):.2f}" == "386.06"
    assert round(P) == 745


def _table_model(device, dtype, rng, num_types=1, knots=200, n=None):
    start = torch.full((num_types * (num_types + 1) // 2,), 0.3, dtype=torch.float64)
    r = 0.3 + np.arange(knots) * (1.2 - 0.3) / (knots - 1)
    lj = 4 * 0.996 * ((0.34 / r) ** 12 - (0.34 / r) ** 6)
    table = np.stack([lj + rng.normal(0, 0.05, knots) for _ in range(start.numel())])
    table[:, -1] = 0.0
    types = torch.tensor(rng.integers(0, num_types, n), dtype=torch.int64, device=device)
    return potential.tabulated(atom_types=types, table=torch.tensor(table, dtype=dtype, device=device),
                               start=start.to(device, dtype), cutoff=1.2), table


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("num_types", [1, 2])
def test_tabulated_vs_oracle(device, dtype, num_types):
    rng = np.random.default_rng(7)
    n = 3000
    cell = np.array([4.2, 3.9, 4.5])
    x = rng.uniform(0, 1, (n, 3)) * cell
    model, table = _table_model(device, dtype, rng, num_types, n=n)
    xd = torch.tensor(x, dtype=dtype, device=device)
    Ld = torch.tensor(cell, dtype=dtype, device=device)
    nl = NeighborList(1.2, skin=0.1)
    nl.allocate(xd, Ld)
    e, F, W = quantity.energy_force_virial(model)(xd, Ld, nl.tensors())
    xw = on.wrap(xd.cpu().numpy(), Ld.cpu().numpy())[0]
    index = nl.tensors().index.cpu().numpy()
    args = (model.atom_types.cpu().numpy(), model.table.detach().cpu().numpy(),
            model.start.cpu().numpy())
    E0, F0, W0 = op.table_efw(xw, Ld.cpu().numpy(), index, *args, 1.2,
                              spacing=model.spacing.cpu().numpy())
    got = (float(e), F.cpu().numpy(), W.cpu().numpy())
    if dtype == torch.float64:
        assert _rel(got[1], F0) <= TOL[dtype] * 3
        assert _rel(got[2], W0) <= TOL[dtype] * 3
        assert abs(got[0] - E0) <= 3 * TOL[dtype] * max(1.0, np.abs(E0))
    else:
        route = table_route(xw, Ld.cpu().numpy(), index, *args, model.spacing.cpu().numpy(), 1.2)
        check_against_route(f"table {num_types} types, fp32", got, route, (E0, F0, W0))


def test_generic_pairs_any_source(device):
    # pairs from elsewhere: shuffled columns, both orientations, padding in between
    rng = np.random.default_rng(3)
    n = 800
    cell = np.array([3.5, 3.5, 3.5])
    x = rng.uniform(0, 3.5, (n, 3))
    xw = on.wrap(x, cell)[0]
    pairs = on.brute_pairs(xw, cell, 1.1).T.copy()
    flip = rng.random(pairs.shape[1]) < 0.5
    pairs[:, flip] = pairs[::-1, flip]
    pad = np.stack([np.arange(500) % n] * 2)
    index = np.concatenate([pairs, pad], axis=1)[:, rng.permutation(pairs.shape[1] + 500)]
    model = potential.lennard_jones(
        atom_types=torch.zeros(n, dtype=torch.int64, device=device),
        sigma=torch.tensor([0.34], dtype=torch.float64, device=device),
        epsilon=torch.tensor([0.996], dtype=torch.float64, device=device), cutoff=1.0)
    from nabladyn_amd.neighbors import NeighborPairs
    pr = NeighborPairs(torch.tensor(index, dtype=torch.int32, device=device), n, 1.1)
    e, F, W = quantity.energy_force_virial(model)(torch.tensor(x, dtype=torch.float64, device=device),
                                                  torch.tensor(cell, dtype=torch.float64, device=device), pr)
    E0, F0, W0 = op.lj_efw(xw, cell, index, np.zeros(n, np.int64), [0.34], [0.996], 1.0)
    assert _rel(F.cpu().numpy(), F0) <= 1e-12
    assert _rel(W.cpu().numpy(), W0) <= 1e-12
    assert abs(float(e) - E0) <= 1e-12 * abs(E0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("kind", ["lj", "table"])
def test_full_rows_equal_the_half_list(device, dtype, kind):
    """quantity over the list's full rows (no atomics) and over its half list (the reference's
    pairs) agree: fp32 to 2e-6 of max|F| (and of W, E), fp64 to 1e-13; the all-pairs path
    (a small box) and the grid path."""
    rng = np.random.default_rng(17)
    for n, length in ((400, 2.6), (6000, 6.0)):
        x = rng.uniform(0, length, (n, 3))
        cell = np.full(3, length)
        if kind == "lj":
            model = potential.lennard_jones(
                atom_types=torch.zeros(n, dtype=torch.int64, device=device),
                sigma=torch.tensor([0.34], dtype=dtype, device=device),
                epsilon=torch.tensor([0.996], dtype=dtype, device=device), cutoff=1.0)
        else:
            model, _ = _table_model(device, dtype, rng, 2, knots=120, n=n)
        xd = torch.tensor(x, dtype=dtype, device=device)
        Ld = torch.tensor(cell, dtype=dtype, device=device)
        nl = NeighborList(model.cutoff, skin=0.1)
        nl.allocate(xd, Ld)
        pairs = nl.tensors()
        assert pairs.full is not None
        efw = quantity.energy_force_virial(model)
        e1, f1, w1 = efw(xd, Ld, pairs)
        e2, f2, w2 = efw(xd, Ld, pairs._replace(full=None))
        tol = 2e-6 if dtype == torch.float32 else 1e-13
        assert float((f1 - f2).abs().max()) <= tol * float(f2.abs().max()), (n, kind)
        assert float((w1 - w2).abs().max()) <= tol * float(w2.abs().max()), (n, kind)
        assert abs(float(e1) - float(e2)) <= tol * max(1.0, abs(float(e2))), (n, kind)
