"""ORACLE — test infrastructure, not product: constant pressure and temperature by isotropic
Martyna-Tobias-Klein dynamics with Nose-Hoover chains, restated in eager torch on the CPU.

Reference files followed:
- ``md/_ensembles/_npt.py:66-266`` the NPT halves: the barostat's and the particles' chains,
  the box momentum's kicks by G = alpha 2K + trace W - 3 P V, the drag exp(-alpha v_eps dt/2),
  the box grown by exp(v_eps dt/2) on both sides of the drift, the conserved energy;
- ``md/_ensembles/_chain.py`` the chain's half step (``oracle.md.chain_half_step``);
- ``md/_ensembles/_verlet.py`` the half kick v + (dt/2) F (1/m);
- ``md/_simulation.py:343-348`` the step order first_half -> neighbor update -> force, virial
  and second_half -> closing;
- the neighbor list's scaled reuse rule (``neighbors/csrc/update/update.cuh:27-116``) through
  ``oracle.md.CPUList``, whose update takes the new cell every step.

Parity status: pinned by invariants in tests/test_oracle_npt.py — with an infinitely heavy box
(tau_b = inf) the dynamics are the oracle's NVT integrator's bit for bit; the conserved energy of
an fp64 run stays flat while the volume moves, and a planted mistake (no drag on the particles)
breaks it; a box started below the target pressure shrinks.
"""

import torch

from . import md as omd

BOLTZMANN = omd.BOLTZMANN


class NPTIntegrator:
    """Isotropic MTK steps in eager torch, in the reference's operation order and dtypes: the
    frame in its dtype; the barostat, the chains and a master cell in float64."""

    def __init__(self, model, x, v, masses, cell, *, dt, skin, temperature, pressure, tau,
                 tau_b, capacity_multiplier=1.25):
        dtype = x.dtype
        self.model, self.dtype = model, dtype
        wide = {"dtype": torch.float64}
        n = x.shape[0]
        self.masses = masses.clone()
        self.inverse_masses = 1.0 / masses[:, None]
        self.cell = cell.clone()
        self.cell64 = cell.double().clone()
        self.x, self.images = omd.wrap(x, cell)
        self.v = v - (masses[:, None] * v).sum(dim=0) / masses.sum()
        self.dt_t = torch.tensor(dt, dtype=dtype)
        self.half_dt_t = torch.tensor(0.5 * dt, dtype=dtype)
        self.dt64 = torch.tensor(dt, **wide)
        self.half_dt64 = torch.tensor(0.5 * dt, **wide)
        self.g = 3 * n - 3
        self.alpha = 1.0 + 1.0 / n
        self.kT = torch.tensor(BOLTZMANN * temperature, **wide)
        self.pressure = torch.tensor(pressure, **wide)
        tb = torch.tensor(tau_b, **wide)
        self.box_mass = 3 * (n + 1) * self.kT * tb.square()
        self.Q = omd.chain_masses(self.g, float(self.kT), float(tau))
        self.Qb = omd.chain_masses(1, float(self.kT), float(tau_b))
        self.chain_x = torch.zeros(3, **wide)
        self.chain_p = torch.zeros(3, **wide)
        self.box_x = torch.zeros(3, **wide)
        self.box_p = torch.zeros(3, **wide)
        self.box_momentum = torch.zeros((), **wide)
        self.list = omd.CPUList(model.cutoff, skin, capacity_multiplier)
        self.list.allocate(self.x.numpy(), self.cell.numpy())
        self.energy, self.forces, self.virial = self.evaluate()
        self.kinetic = self.kinetic_energy().double()
        self.rebuild_steps = []
        self.steps = 0

    def evaluate(self):
        strain = self.x.new_zeros((3, 3))
        g = torch.func.grad_and_value(self.model.energy, argnums=(0, 1))
        (gx, gs), e = g(self.x, strain, self.cell, torch.as_tensor(self.list.index).long())
        return e, -gx, -gs

    def kinetic_energy(self):
        return 0.5 * (self.masses[:, None] * self.v.square()).sum()

    def box_force(self):
        """G = alpha 2K + trace W - 3 P V with the carried K (``_npt.py:232-238``)."""
        return (self.alpha * 2.0 * self.kinetic + self.virial.trace().double()
                - 3.0 * self.pressure * self.cell64.prod())

    def chain(self):
        factor, self.chain_x, self.chain_p = omd.chain_half_step(
            self.kinetic, self.chain_x, self.chain_p, self.Q, self.g, self.kT, self.dt64)
        self.v = self.v * factor.to(self.dtype)
        self.kinetic = self.kinetic * factor.square()

    def box_chain(self):
        kinetic = self.box_momentum.square() / (2.0 * self.box_mass)
        factor, self.box_x, self.box_p = omd.chain_half_step(
            kinetic, self.box_x, self.box_p, self.Qb, 1, self.kT, self.dt64)
        self.box_momentum = self.box_momentum * factor

    def drag(self):
        rate = self.box_momentum / self.box_mass
        return torch.exp(-self.alpha * rate * self.half_dt64).to(self.dtype), rate

    def step(self):
        # first half (_npt.py:154-174)
        self.box_chain()
        self.chain()
        self.box_momentum = self.box_momentum + self.half_dt64 * self.box_force()
        drag, rate = self.drag()
        self.v = (self.v * drag) + self.half_dt_t * self.forces * self.inverse_masses
        growth = torch.exp(rate * self.half_dt64)
        scale = growth.to(self.dtype)
        moved = (self.x * scale + self.dt_t * self.v) * scale
        self.cell64 = self.cell64 * growth * growth
        self.cell = self.cell64.to(self.dtype)
        self.x, crossings = omd.wrap(moved, self.cell)
        self.images = self.images + crossings
        # the list, at the new cell
        self.steps += 1
        if self.list.update(self.x.numpy(), self.cell.numpy()):
            self.rebuild_steps.append(self.steps)
        self.energy, self.forces, self.virial = self.evaluate()
        # second half (_npt.py:176-186)
        drag, _ = self.drag()
        self.v = (self.v + self.half_dt_t * self.forces * self.inverse_masses) * drag
        self.kinetic = self.kinetic_energy().double()
        # closing (_npt.py:188-191)
        self.box_momentum = self.box_momentum + self.half_dt64 * self.box_force()
        self.chain()
        self.box_chain()

    def volume(self) -> float:
        return float(self.cell64.prod())

    def conserved_energy(self) -> float:
        """U + K + both chains' energies + P V + p_eps^2 / (2 M) (``_npt.py:205-230``)."""
        baths = (omd.heat_bath_energy(self.chain_x, self.chain_p, self.Q, self.g, self.kT)
                 + omd.heat_bath_energy(self.box_x, self.box_p, self.Qb, 1, self.kT))
        return float(self.energy.double() + self.kinetic_energy().double() + baths
                     + self.pressure * self.cell64.prod()
                     + self.box_momentum.square() / (2.0 * self.box_mass))


def pressure_of(kinetic: float, virial_trace: float, volume: float) -> float:
    """P = (2K + trace W) / (3V) (``observables.py:18-41``), kJ/mol/nm^3."""
    return (2.0 * kinetic + virial_trace) / (3.0 * volume)


__all__ = ("NPTIntegrator", "pressure_of")
