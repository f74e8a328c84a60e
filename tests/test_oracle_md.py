"""Pins the MD oracle: maxwell_boltzmann's exact temperature and zero momentum, the chain half
step's time reversibility, NVE energy conservation and the NVT conserved energy on a small
fp64 system (md/_velocities.py, md/_ensembles/_chain.py, _nve.py, _nvt.py)."""

import numpy as np
import torch

from oracle import md as omd
from oracle import neighbors as on


def test_maxwell_boltzmann_exact_temperature():
    masses = torch.tensor(np.random.default_rng(0).uniform(10, 50, 300))
    v = omd.maxwell_boltzmann(masses, 150.0, 7)
    momentum = (masses[:, None] * v).sum(dim=0)
    assert float(momentum.abs().max()) < 1e-12
    kinetic = 0.5 * float((masses[:, None] * v.square()).sum())
    temperature = 2 * kinetic / ((3 * 300 - 3) * omd.BOLTZMANN)
    assert abs(temperature - 150.0) < 1e-9
    assert torch.equal(v, omd.maxwell_boltzmann(masses, 150.0, 7))


def test_chain_half_step_is_reversible():
    g, kT, dt = 1533, omd.BOLTZMANN * 120.0, 0.004
    masses = omd.chain_masses(g, kT, 0.1)
    q0 = torch.tensor([0.01, -0.02, 0.005], dtype=torch.float64)
    p0 = torch.tensor([3.0, -1.5, 0.7], dtype=torch.float64)
    k0 = torch.tensor(700.0, dtype=torch.float64)
    factor, q1, p1 = omd.chain_half_step(k0, q0, p0, masses, g, kT, dt)
    k1 = k0 * factor.square()
    back, q2, p2 = omd.chain_half_step(k1, q1, p1, masses, g, kT, -dt)
    assert float((q2 - q0).abs().max()) < 1e-12
    assert float((p2 - p0).abs().max()) < 1e-10
    assert abs(float(factor * back) - 1.0) < 1e-12


def _system(dtype=torch.float64):
    x = torch.tensor(on.jittered_lattice(6, 2.4, 0.02, seed=3), dtype=dtype)
    masses = torch.full((216,), 39.948, dtype=dtype)
    v = omd.maxwell_boltzmann(masses, 100.0, 4)
    model = omd.LJModel(np.zeros(216, np.int64), [0.34], [0.996], 1.0, dtype)
    return model, x, v, masses, torch.full((3,), 2.4, dtype=dtype)


def test_nve_conserves_energy():
    model, x, v, m, L = _system()
    sim = omd.Integrator(model, x, v, m, L, dt=0.002, skin=0.1)
    e0 = sim.total_energy()
    energies = []
    for _ in range(200):
        sim.step()
        energies.append(sim.total_energy())
    assert np.abs(np.array(energies) - e0).max() / 216 < 2e-4
    assert sim.list.rebuilds > 1


def test_nvt_conserved_energy_and_momentum():
    model, x, v, m, L = _system()
    sim = omd.Integrator(model, x, v, m, L, dt=0.002, skin=0.1, temperature=100.0, tau=0.1)
    c0 = sim.conserved_energy()
    values = []
    for _ in range(200):
        sim.step()
        values.append(sim.conserved_energy())
    assert np.abs(np.array(values) - c0).max() / 216 < 3e-4
    assert float((m[:, None] * sim.v).sum(dim=0).abs().max()) < 1e-10


def test_sum_model_on_one_list_equals_each_part_on_its_own_list():
    """The reference hands each model of a sum the list of its own cutoff
    (potential/_sum/_potential.py:34-43,86-99); the oracle's SumModel masks each part at its
    cutoff on the longest list. Both give the same energy and forces in fp64 (two types, LJ at
    0.8 nm plus a table at 1.0 nm); a part left unmasked would not."""
    rng = np.random.default_rng(12)
    cell = np.full(3, 2.5)
    x = on.wrap(on.jittered_lattice(7, 2.5, 0.03, seed=12), cell)[0]
    types = rng.integers(0, 2, len(x))
    r = 0.3 + 0.7 / 49 * np.arange(50)
    table = np.stack([4.0 * ((0.33 / r) ** 12 - (0.33 / r) ** 6) + rng.normal(0, 0.05, 50)
                      for _ in range(3)]).clip(None, 40.0)
    table[:, -1] = 0.0
    lj = omd.LJModel(types, [0.30, 0.32, 0.34], [0.8, 1.0, 1.2], 0.8, torch.float64)
    tab = omd.TableModel(types, table, [0.3, 0.3, 0.3], 1.0, torch.float64)
    total = omd.SumModel(lj, tab)
    assert total.cutoff == 1.0
    xt, ct = torch.tensor(x), torch.tensor(cell)
    strain = torch.zeros((3, 3), dtype=torch.float64)
    far = torch.tensor(on.brute_pairs(x, cell, 1.0).T).long()
    near = torch.tensor(on.brute_pairs(x, cell, 0.8).T).long()

    def efw(energy, index):
        g = torch.func.grad_and_value(energy, argnums=0)
        gx, e = g(xt, strain, ct, index)
        return float(e), gx

    e_sum, g_sum = efw(total.energy, far)
    e_lj, g_lj = efw(lj.energy, near)
    e_tab, g_tab = efw(tab.energy, far)
    assert abs(e_sum - (e_lj + e_tab)) <= 1e-12 * abs(e_sum)
    assert float((g_sum - (g_lj + g_tab)).abs().max()) <= 1e-12 * float(g_sum.abs().max())
    # unmasked, the LJ part would count the pairs between 0.8 and 1.0 nm as well
    unmasked = omd.LJModel(types, [0.30, 0.32, 0.34], [0.8, 1.0, 1.2], 1.0, torch.float64)
    e_wide, _ = efw(unmasked.energy, far)
    assert abs(e_wide - e_lj) > 1e-3 * abs(e_lj)
