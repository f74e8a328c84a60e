"""NPT (isotropic MTK with Nose-Hoover chains, ``md/_ensembles/_npt.py``) on the fused MD step,
against the oracle's eager restatement (``oracle/npt.py``).

- fp64, 512 argon atoms, 80 steps: positions, velocities, cell, box momentum and both chains
  within 1e-9 relative of the oracle's, and the list rebuilt at the same steps (the reference's
  reuse rule scaled with the changing cell, ``update.cuh:27-75``).
- fp32, 1,500 steps at 4 fs: the conserved energy's spread stays small while the volume moves.
- A run continued from ``state()`` (with its barostat) follows the original.
- A box compressed past the grid's margin stops the run with ``ValueError``.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import md, potential, units
from oracle import md as omd
from oracle import neighbors as on
from oracle.npt import NPTIntegrator

pytestmark = pytest.mark.gpu

MASS = 39.948


def _lj(n, dtype, device):
    return potential.lennard_jones(
        atom_types=torch.zeros(n, dtype=torch.int64, device=device),
        sigma=torch.tensor([0.34], dtype=dtype, device=device),
        epsilon=torch.tensor([0.996], dtype=dtype, device=device), cutoff=1.0)


def _argon(dtype, length=2.90, seed=0):
    x = torch.tensor(on.jittered_lattice(8, length, 0.02, seed=seed), dtype=dtype)
    masses = torch.full((512,), MASS, dtype=dtype)
    v = md.maxwell_boltzmann(masses, 120.0, 1)
    return x, v, masses, torch.full((3,), length, dtype=dtype)


KW = dict(timestep=0.004, temperature=120.0, pressure=400 * units.BAR, thermostat_timescale=0.1,
          barostat_timescale=0.5)


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_npt_matches_the_oracle_fp64(device):
    dtype = torch.float64
    x, v, m, L = _argon(dtype)
    sim = md.npt(_lj(512, dtype, device), x.to(device), v.to(device), m.to(device), L.to(device),
                 **KW)
    ref = NPTIntegrator(omd.LJModel(np.zeros(512, np.int64), [0.34], [0.996], 1.0, dtype),
                        x, v, m, L, dt=0.004, skin=0.1, temperature=120.0,
                        pressure=400 * units.BAR, tau=0.1, tau_b=0.5)
    rebuilt = []
    for step in range(1, 81):
        ref.step()
        count = sim._neighbors._rebuild_count
        sim._launch_step()
        sim._step += 1
        if sim._neighbors._rebuild_count != count:
            rebuilt.append(step)
    sim._check(sim._step)
    state = sim.state()
    unwrapped = state.positions.cpu() + state.images.cpu().to(dtype) * state.cell.cpu()
    assert _rel(unwrapped, ref.x + ref.images.to(dtype) * ref.cell) <= 1e-9
    assert _rel(state.velocities.cpu(), ref.v) <= 1e-9
    assert _rel(state.cell.cpu(), ref.cell) <= 1e-12
    assert abs(float(state.barostat.momentum) - float(ref.box_momentum)) <= \
        1e-9 * abs(float(ref.box_momentum))
    assert torch.allclose(state.thermostat.momenta.cpu(), ref.chain_p, rtol=1e-8, atol=1e-12)
    assert torch.allclose(state.barostat.chain.momenta.cpu(), ref.box_p, rtol=1e-8, atol=1e-12)
    assert abs(float(state.cell[0]) - 2.90) > 1e-5  # the box moved
    assert len(ref.rebuild_steps) >= 2
    assert rebuilt == ref.rebuild_steps
    panel = sim.sample()
    assert abs(panel["conserved_energy"] - ref.conserved_energy()) <= 1e-8 * abs(ref.conserved_energy())
    assert abs(panel["volume"] - ref.volume()) <= 1e-12 * ref.volume()


def test_npt_conserved_energy_fp32(device):
    dtype = torch.float32
    x, v, m, L = _argon(dtype)
    sim = md.npt(_lj(512, dtype, device), x.to(device), v.to(device), m.to(device), L.to(device),
                 **KW)
    samples = sim.run(1500, sample_every=50)
    conserved = samples["conserved_energy"].numpy()
    volume = samples["volume"].numpy()
    assert np.ptp(volume) > 0.05, np.ptp(volume)
    assert np.ptp(conserved) < 1.0, np.ptp(conserved)  # kJ/mol over 6 ps, 512 atoms


def test_npt_state_continues_run(device):
    dtype = torch.float64
    x, v, m, L = _argon(dtype)
    model = _lj(512, dtype, device)
    a = md.npt(model, x.to(device), v.to(device), m.to(device), L.to(device), **KW)
    a.run(30)
    s = a.state()
    b = md.npt(model, s.positions, s.velocities, m.to(device), s.cell, images=s.images,
               step=s.step, thermostat=s.thermostat, barostat=s.barostat, **KW)
    a.run(10)
    b.run(10)
    sa, sb = a.state(), b.state()
    assert float((sa.positions - sb.positions).abs().max()) < 1e-9
    assert float((sa.cell - sb.cell).abs().max()) < 1e-12
    assert sa.step == sb.step == 40


def test_npt_box_compressed_past_the_grid_stops(device):
    """A dilute box (density 10 nm^-3) driven at 5,000 bar collapses; once it has shrunk past
    the sub-bin grid's margin the run stops with ValueError."""
    dtype = torch.float32
    x, v, m, L = _argon(dtype, length=3.7)
    sim = md.npt(_lj(512, dtype, device), x.to(device), v.to(device), m.to(device), L.to(device),
                 timestep=0.004, temperature=120.0, pressure=5000 * units.BAR,
                 thermostat_timescale=0.1, barostat_timescale=0.2, capacity_multiplier=2.0)
    with pytest.raises(ValueError, match="shrank|finite|capacity"):
        sim.run(3000)
