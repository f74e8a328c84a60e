"""A sum of two models in the MD step (``potential.sum``, ``potential/_sum/_potential.py:86-99``)
against the oracle's eager integrator with ``SumModel``, which hands each model the pairs within
its own cutoff as the reference does.

- fp64 NVT, 512 argon atoms, LJ at 0.8 nm plus a one-type 200-knot table at 1.0 nm, 20 steps:
  positions within 1e-9 (relative, max norm), velocities and forces within 1e-8, the chain's
  state within 1e-8, and the list rebuilt at the same steps.
- The first frame of a two-type sum (LJ at 0.9 nm plus a table at 1.2 nm), fp64 within 3e-12 of
  the oracle's closed forms added; fp32 forces within 1e-5 of the fp64 values.
- Sums the step does not take are refused: three models, models of other atom types, a sum in
  a SimulationBatch.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import md, potential
from oracle import md as omd
from oracle import neighbors as on
from oracle import potentials as op

pytestmark = pytest.mark.gpu

MASS = 39.948


def _table_values(knots, start, cutoff, rows, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    r = start + (cutoff - start) / (knots - 1) * np.arange(knots)
    shape = scale * 4 * 0.996 * ((0.36 / r) ** 12 - (0.36 / r) ** 6)
    table = np.stack([np.clip(shape, None, 30.0) + rng.normal(0, 0.02, knots)
                      for _ in range(rows)])
    table[:, -1] = 0.0
    return table


def _sum(types, dtype, device, lj_cutoff, table_cutoff, table):
    rows = table.shape[0]
    t = torch.tensor(types, dtype=torch.int64, device=device)
    nt = {1: 1, 3: 2}[rows]
    lj = potential.lennard_jones(
        atom_types=t, sigma=torch.tensor([0.34, 0.33, 0.35][:rows], dtype=dtype, device=device),
        epsilon=torch.tensor([0.996, 0.9, 1.1][:rows], dtype=dtype, device=device),
        cutoff=lj_cutoff)
    tab = potential.tabulated(
        atom_types=t, table=torch.tensor(table, dtype=dtype, device=device),
        start=torch.full((rows,), 0.3, dtype=dtype, device=device), cutoff=table_cutoff)
    assert lj._num_types == nt
    return potential.sum(lj, tab)


def test_sum_nvt_trajectory_matches_oracle_fp64(device):
    dtype = torch.float64
    x = torch.tensor(on.jittered_lattice(8, 2.90, 0.02, seed=0), dtype=dtype)
    masses = torch.full((512,), MASS, dtype=dtype)
    v = md.maxwell_boltzmann(masses, 120.0, 1)
    cell = torch.full((3,), 2.90, dtype=dtype)
    types = np.zeros(512, np.int64)
    table = _table_values(200, 0.3, 1.0, 1, seed=3)
    model = _sum(types, dtype, device, 0.8, 1.0, table)
    sim = md.nvt(model, x.to(device), v.to(device), masses.to(device), cell.to(device),
                 timestep=0.004, temperature=120.0, thermostat_timescale=0.1, skin=0.1)
    ref = omd.Integrator(
        omd.SumModel(omd.LJModel(types, [0.34], [0.996], 0.8, dtype),
                     omd.TableModel(types, table, [0.3], 1.0, dtype)),
        x, v, masses, cell, dt=0.004, skin=0.1, temperature=120.0, tau=0.1)
    rebuilt, expected = [], []
    for step in range(1, 21):
        before = ref.list.rebuilds
        ref.step()
        if ref.list.rebuilds != before:
            expected.append(step)
        count = sim._neighbors._rebuild_count
        sim._launch_step()
        sim._step += 1
        if sim._neighbors._rebuild_count != count:
            rebuilt.append(step)
    sim._check(sim._step)
    state = sim.state()
    pos = state.positions.cpu() + state.images.cpu().to(dtype) * cell
    ref_pos = ref.x + ref.images.to(dtype) * cell
    assert float((pos - ref_pos).abs().max() / ref_pos.abs().max()) <= 1e-9
    assert float((state.velocities.cpu() - ref.v).abs().max() / ref.v.abs().max()) <= 1e-8
    assert float((state.forces.cpu() - ref.forces).abs().max() / ref.forces.abs().max()) <= 1e-8
    assert torch.allclose(state.thermostat.momenta.cpu(), ref.chain_p, rtol=1e-8, atol=1e-12)
    assert rebuilt == expected


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_sum_first_frame_two_types(device, dtype):
    rng = np.random.default_rng(8)
    cell = np.array([4.4, 4.1, 4.6])
    x = on.wrap(on.jittered_lattice(11, 4.4, 0.03, seed=8) * (cell / 4.4), cell)[0]
    n = len(x)
    types = rng.integers(0, 2, n)
    table = _table_values(120, 0.3, 1.2, 3, seed=9)
    model = _sum(types, dtype, device, 0.9, 1.2, table)
    xd = torch.tensor(x, dtype=dtype, device=device)
    sim = md.nve(model, xd, torch.zeros_like(xd), torch.full((n,), MASS, dtype=dtype,
                 device=device), torch.tensor(cell, dtype=dtype, device=device),
                 timestep=0.002, skin=0.2)
    xw = on.wrap(xd.double().cpu().numpy(), cell)[0]
    far = on.brute_pairs(xw, cell, 1.2).T.copy()
    near = on.brute_pairs(xw, cell, 0.9).T.copy()
    e1, f1, w1 = op.lj_efw(xw, cell, near, types, [0.34, 0.33, 0.35], [0.996, 0.9, 1.1], 0.9)
    table64 = torch.tensor(table, dtype=dtype).double().numpy()
    spacing = model.parts[1].spacing.double().cpu().numpy()
    e2, f2, w2 = op.table_efw(xw, cell, far, types, table64, np.full(3, 0.3), 1.2,
                              spacing=spacing)
    E0, F0, W0 = e1 + e2, f1 + f2, w1 + w2
    F = sim._to_atoms(sim.forces).double().cpu().numpy()
    W = sim.virial.double().cpu().numpy()
    tol = 3e-12 if dtype == torch.float64 else 1e-5
    assert np.abs(F - F0).max() <= tol * np.abs(F0).max()
    assert np.abs(W - W0).max() <= tol * np.abs(W0).max()
    assert abs(float(sim.energy) - E0) <= tol * max(1.0, abs(E0))


def test_sums_the_step_does_not_take(device):
    dtype = torch.float64
    x = torch.tensor(on.jittered_lattice(8, 2.90, 0.02, seed=0), dtype=dtype, device=device)
    m = torch.full((512,), MASS, dtype=dtype, device=device)
    cell = torch.full((3,), 2.90, dtype=dtype, device=device)
    zeros = np.zeros(512, np.int64)
    table = _table_values(50, 0.3, 1.0, 1, seed=3)
    one = _sum(zeros, dtype, device, 0.8, 1.0, table)
    three = potential.sum(one.parts[0], one.parts[1], one.parts[0])
    with pytest.raises(TypeError, match="exactly two"):
        md.nve(three, x, torch.zeros_like(x), m, cell, timestep=0.002)
    other = potential.lennard_jones(
        atom_types=torch.tensor(np.arange(512) % 2, device=device),
        sigma=torch.full((3,), 0.34, dtype=dtype, device=device),
        epsilon=torch.full((3,), 0.9, dtype=dtype, device=device), cutoff=0.8)
    with pytest.raises(TypeError, match="same atom types"):
        md.nve(potential.sum(other, one.parts[1]), x, torch.zeros_like(x), m, cell,
               timestep=0.002)
    sim = md.Simulation(one, x, torch.zeros_like(x), m, cell, timestep=0.002, skin=0.1,
                        capacity_multiplier=1.25, images=None, step=0, graphs=False)
    with pytest.raises(ValueError, match="stepped alone"):
        md.SimulationBatch([sim])
