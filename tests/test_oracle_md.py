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
