"""``potential.sum`` on the drop-in path: a short-range LJ (rc 0.9 nm) plus a 200-knot table
(rc 1.2 nm) on 2,000 random atoms, evaluated by ``quantity.energy_force_virial``.

- fp64: against the oracle's closed forms of the parts, each on brute-force pairs within its
  own cutoff, to 3e-12. fp32: against the reference's own fp32 route of each part (autograd of
  its energy, restated by the oracle) on the same pairs, added: max|dF| / max|F|,
  max|dW| / max|W| and |dE| / max(1, |E|) <= 1e-5, with the fp64 errors printed beside.
- A list per cutoff changes the cost, never the result (``potential/_sum/_potential.py:94-95``):
  the sum on a list with an inner list at 0.9 nm equals the sum on the 1.2 nm list alone, and
  the parts evaluated alone add up to the sum: to 1e-6 of max|F| (and of W, E) in fp32, 1e-12
  in fp64 (the half-list kernel adds forces with atomics, so two evaluations may round in
  another order).
- The checks of ``potential.sum`` (TypeError, ValueError) and MD's refusal of a sum.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import md, potential, quantity
from nabladyn_amd.neighbors import NeighborList
from oracle import neighbors as on
from oracle import potentials as op
from _parity import check_against_route, lj_route, table_route

pytestmark = pytest.mark.gpu


def _system(dtype, device, seed=21, n=2000):
    rng = np.random.default_rng(seed)
    cell = np.array([4.6, 4.3, 5.0])
    x = rng.uniform(0, 1, (n, 3)) * cell
    types = rng.integers(0, 2, n)
    r = 0.3 + np.arange(200) * (1.2 - 0.3) / 199
    # one row per unordered type pair (3 for 2 types), last knot 0
    table = np.stack([np.clip(4 * e * ((0.32 / r) ** 12 - (0.32 / r) ** 6), None, 40.0)
                      + rng.normal(0, 0.02, 200) for e in (0.8, 0.6, 0.5)])
    table[:, -1] = 0.0
    t = torch.tensor(types, dtype=torch.int64, device=device)
    lj = potential.lennard_jones(atom_types=t, sigma=torch.tensor([0.34, 0.30, 0.28], dtype=dtype, device=device),
                                 epsilon=torch.tensor([0.9, 0.6, 0.4], dtype=dtype, device=device), cutoff=0.9)
    tab = potential.tabulated(atom_types=t, table=torch.tensor(table, dtype=dtype, device=device),
                              start=torch.full((3,), 0.3, dtype=dtype, device=device), cutoff=1.2)
    return x, cell, types, table, lj, tab


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_sum_vs_oracle(device, dtype):
    x, cell, types, table, lj, tab = _system(dtype, device)
    model = potential.sum(lj, tab)
    assert model.cutoff == 1.2 and model._list_cutoffs == (0.9, 1.2)
    xd = torch.tensor(x, dtype=dtype, device=device)
    Ld = torch.tensor(cell, dtype=dtype, device=device)
    nl = NeighborList(1.2, skin=0.0, inner_cutoffs=(0.9,))
    nl.allocate(xd, Ld)
    e, F, W = quantity.energy_force_virial(model)(xd, Ld, nl.tensors())
    xw = on.wrap(xd.cpu().numpy(), cell)[0]
    near = on.brute_pairs(xw, cell, 0.9).T.copy()
    far = on.brute_pairs(xw, cell, 1.2).T.copy()
    f32 = lambda a: np.asarray(a, dtype=xd.cpu().numpy().dtype).astype(np.float64)
    E1, F1, W1 = op.lj_efw(xw, cell, near, types, f32([0.34, 0.30, 0.28]), f32([0.9, 0.6, 0.4]), 0.9)
    E2, F2, W2 = op.table_efw(xw, cell, far, types, f32(table), f32([0.3] * 3), 1.2,
                              spacing=tab.spacing.cpu().numpy())
    F0, W0, E0 = F1 + F2, W1 + W2, E1 + E2
    if dtype == torch.float64:
        tol = 3e-12
        assert np.abs(F.cpu().double().numpy() - F0).max() <= tol * np.abs(F0).max()
        assert np.abs(W.cpu().double().numpy() - W0).max() <= tol * np.abs(W0).max()
        assert abs(float(e) - E0) <= tol * max(1.0, abs(E0))
        return
    # fp32: the reference's route, each part on its own pairs, added as its Sum adds them
    native = lambda a: np.asarray(a, dtype=np.float32)
    r1 = lj_route(xw, cell.astype(np.float32), near, types, native([0.34, 0.30, 0.28]),
                  native([0.9, 0.6, 0.4]), 0.9)
    r2 = table_route(xw, cell.astype(np.float32), far, types, native(table), native([0.3] * 3),
                     tab.spacing.cpu().numpy(), 1.2)
    route = (float(np.float32(r1[0]) + np.float32(r2[0])), r1[1] + r2[1], r1[2] + r2[2])
    check_against_route("sum LJ 0.9 + table 1.2, fp32", (float(e), F.cpu().numpy(), W.cpu().numpy()),
                        route, (E0, F0, W0))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_lists_change_the_cost_not_the_result(device, dtype):
    x, cell, _, _, lj, tab = _system(dtype, device, seed=22)
    model = potential.sum(lj, tab)
    xd = torch.tensor(x, dtype=dtype, device=device)
    Ld = torch.tensor(cell, dtype=dtype, device=device)
    with_inner = NeighborList(1.2, skin=0.1, inner_cutoffs=(0.9,))
    with_inner.allocate(xd, Ld)
    alone = NeighborList(1.2, skin=0.1)
    alone.allocate(xd, Ld)
    efw = quantity.energy_force_virial(model)
    e1, F1, W1 = efw(xd, Ld, with_inner.tensors())
    e2, F2, W2 = efw(xd, Ld, alone.tensors())
    # the half-list kernel adds forces with atomics: two evaluations may round differently
    tol = 1e-6 if dtype == torch.float32 else 1e-12
    assert float((F1 - F2).abs().max()) <= tol * float(F1.abs().max())
    assert float((W1 - W2).abs().max()) <= tol * float(W1.abs().max())
    assert abs(float(e1 - e2)) <= tol * abs(float(e1))
    # the parts alone, each on its own list, add up to the sum
    inner = with_inner.tensors().inner[0]
    el, Fl, Wl = quantity.energy_force_virial(lj)(xd, Ld, inner)
    et, Ft, Wt = quantity.energy_force_virial(tab)(xd, Ld, with_inner.tensors())
    assert float((F1 - (Fl + Ft)).abs().max()) <= tol * float(F1.abs().max())
    assert float((W1 - (Wl + Wt)).abs().max()) <= tol * float(W1.abs().max())
    assert abs(float(e1 - (el + et))) <= tol * abs(float(e1))


def test_sum_checks(device):
    x, cell, _, _, lj, tab = _system(torch.float32, device)
    with pytest.raises(ValueError):
        potential.sum(lj)
    with pytest.raises(TypeError):
        potential.sum(lj, "table")
    other = potential.lennard_jones(atom_types=torch.zeros(5, dtype=torch.int64, device=device),
                                    sigma=torch.tensor([0.34], device=device),
                                    epsilon=torch.tensor([1.0], device=device), cutoff=0.9)
    with pytest.raises(ValueError):
        potential.sum(lj, other)
    model = potential.sum(lj, tab)
    assert set(name.split(".")[1] for name, _ in model.named_parameters()) <= {"0", "1"}
    xd = torch.tensor(x, dtype=torch.float32, device=device)
    masses = torch.full((xd.shape[0],), 39.948, device=device)
    # the MD step takes a sum of two models (tests/test_gpu_md_sum.py holds its parity)
    # (random positions overlap: the first frame only, no steps)
    sim = md.nve(model, xd, torch.zeros_like(xd), masses,
                 torch.tensor(cell, dtype=torch.float32, device=device), timestep=0.002)
    assert bool(torch.isfinite(sim.energy))
