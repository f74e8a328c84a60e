"""The oracle's NPT integrator (isotropic MTK with Nose-Hoover chains, ``md/_ensembles/_npt.py``)
pinned by what the mathematics fixes, on the CPU:

- an infinitely heavy box (tau_b = inf) has no rate, so the dynamics are the oracle's NVT
  integrator's, bit for bit;
- the MTK conserved energy (U + K + both chains + P V + p_eps^2 / 2M) stays flat in fp64,
  while the volume moves: a planted mistake (the particles' drag left out) breaks it;
- the box responds to the pressure difference with the right sign: compressed below the target
  pressure, the box shrinks.
"""

import math

import numpy as np
import torch

from oracle import md as omd
from oracle import neighbors as on
from oracle.npt import NPTIntegrator, pressure_of

MASS = 39.948
BAR = 0.0602214076  # kJ/mol/nm^3 per bar


def _argon(dtype=torch.float64, seed=0, length=2.90):
    x = torch.tensor(on.jittered_lattice(8, length, 0.02, seed=seed), dtype=dtype)
    masses = torch.full((512,), MASS, dtype=dtype)
    v = omd.maxwell_boltzmann(masses, 120.0, 1)
    return x, v, masses, torch.full((3,), length, dtype=dtype)


def _model(dtype=torch.float64):
    return omd.LJModel(np.zeros(512, np.int64), [0.34], [0.996], 1.0, dtype)


def test_infinite_box_mass_is_nvt_bit_for_bit():
    x, v, m, L = _argon()
    npt = NPTIntegrator(_model(), x, v, m, L, dt=0.004, skin=0.1, temperature=120.0,
                        pressure=1.0, tau=0.1, tau_b=math.inf)
    nvt = omd.Integrator(_model(), x, v, m, L, dt=0.004, skin=0.1, temperature=120.0, tau=0.1)
    for _ in range(15):
        npt.step()
        nvt.step()
    assert torch.equal(npt.x, nvt.x)
    assert torch.equal(npt.v, nvt.v)
    assert torch.equal(npt.cell, L)
    assert torch.equal(npt.chain_p, nvt.chain_p)


def _conserved_spread(integrator, steps):
    values, volumes = [], []
    for _ in range(steps):
        integrator.step()
        values.append(integrator.conserved_energy())
        volumes.append(integrator.volume())
    return np.ptp(values), np.ptp(volumes)


def test_conserved_energy_is_flat_while_the_box_moves():
    x, v, m, L = _argon()
    kw = dict(dt=0.002, skin=0.1, temperature=120.0, pressure=400 * BAR, tau=0.1, tau_b=0.5)
    spread, volumes = _conserved_spread(NPTIntegrator(_model(), x, v, m, L, **kw), 150)
    assert volumes > 0.01, volumes  # the box moves (nm^3)
    assert spread < 0.05, spread    # kJ/mol, 512 atoms

    class NoDrag(NPTIntegrator):  # a planted mistake: the particles feel no box drag
        def drag(self):
            _, rate = super().drag()
            return torch.ones((), dtype=self.dtype), rate

    broken, _ = _conserved_spread(NoDrag(_model(), x, v, m, L, **kw), 150)
    assert broken > 5 * max(spread, 1e-3), (broken, spread)


def test_box_shrinks_below_the_target_pressure():
    """Start at a low density (box 3.1 nm, well below the pressure asked): the box momentum
    turns negative and the volume falls."""
    x, v, m, L = _argon(length=3.1)
    npt = NPTIntegrator(_model(), x, v, m, L, dt=0.004, skin=0.1, temperature=120.0,
                        pressure=2000 * BAR, tau=0.1, tau_b=0.3, capacity_multiplier=2.0)
    start = npt.volume()
    p0 = pressure_of(float(npt.kinetic), float(npt.virial.trace()), start)
    assert p0 < 2000 * BAR
    for _ in range(40):
        npt.step()
    assert float(npt.box_momentum) < 0
    assert npt.volume() < start
