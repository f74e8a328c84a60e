"""ORACLE — test infrastructure, not product: velocity Verlet, the Nose-Hoover chain, and the
initial velocities, restated on the CPU in torch (eager, no compile).

Reference files followed:
- ``md/_ensembles/_verlet.py:8-31`` half_kick, drift;
- ``md/_ensembles/_nve.py:39-58`` NVE halves;
- ``md/_ensembles/_nvt.py:41-129`` NVT halves, fp64 chain, K carried between programs;
- ``md/_ensembles/_chain.py:23-105`` chain masses, half step, heat-bath energy;
- ``md/_velocities.py:42-65`` maxwell_boltzmann (torch CPU generator, fp64, COM removed,
  exact T over 3N - 3);
- ``md/_simulation.py:137-196,316-364`` first frame (wrap, COM removal, first force) and the
  step order first_half -> neighbor update -> force + second_half -> closing;
- ``observables.py:18-41`` K, T, P;
- ``potential/_sum/_potential.py:86-99`` a sum of models (``SumModel``), each on the pairs within
  its own cutoff; ``TableModel`` evaluates ``oracle.potentials.t_table_energy``.

Parity status: pinned by invariants. Energy conservation of NVE (drift), zero COM momentum,
``maxwell_boltzmann`` giving exactly T, the conserved energy of NVT, and the chain's
time-reversibility (a half step and its reverse return the start to rounding) are checked in
tests/test_oracle_md.py. No reference trajectory exists to compare with. ``SumModel`` on one
list equals the parts evaluated each on the list of its own cutoff, as the reference hands them
out (tests/test_oracle_md.py).
"""

import math

import numpy as np
import torch

from . import neighbors as on
from . import potentials as op

BOLTZMANN = 0.00831446261815324  # kJ/(mol K) (units.py)
BAR = 0.0602214076  # kJ/(mol nm^3)


def maxwell_boltzmann(masses: torch.Tensor, temperature: float, seed: int) -> torch.Tensor:
    """``md/_velocities.py:42-65`` restated."""
    generator = torch.Generator().manual_seed(int(seed))
    host = masses.detach().to("cpu", torch.float64)
    v = torch.randn((host.numel(), 3), generator=generator, dtype=torch.float64)
    v *= torch.sqrt(BOLTZMANN * float(temperature) / host)[:, None]
    v -= (host[:, None] * v).sum(dim=0) / host.sum()
    kinetic = 0.5 * float((host[:, None] * v.square()).sum())
    target = 0.5 * (3 * host.numel() - 3) * BOLTZMANN * float(temperature)
    if kinetic > 0:
        v *= math.sqrt(target / kinetic)
    return v.to(masses.device, masses.dtype)


def wrap(positions: torch.Tensor, cell: torch.Tensor):
    """``space.py:22-38`` in torch."""
    crossings = torch.floor(positions / cell)
    wrapped = positions - crossings * cell
    below = wrapped < 0
    wrapped = torch.where(below, wrapped + cell, wrapped)
    crossings = torch.where(below, crossings - 1, crossings)
    above = wrapped >= cell
    wrapped = torch.where(above, wrapped - cell, wrapped)
    crossings = torch.where(above, crossings + 1, crossings)
    return wrapped, crossings.to(torch.int64)


def chain_masses(g: int, kT: float, tau: float) -> torch.Tensor:
    """``_chain.py:23-26``."""
    kT = torch.tensor(kT, dtype=torch.float64)
    base = kT * torch.tensor(tau, dtype=torch.float64).square()
    return torch.stack((g * base, base, base))


def chain_half_step(kinetic, positions, momenta, masses, g, kT, dt):
    """``_chain.py:29-76`` restated (fp64 tensors). Returns (factor, positions, momenta)."""
    kT = torch.as_tensor(kT, dtype=torch.float64)
    dt = torch.as_tensor(dt, dtype=torch.float64)
    p, mass = list(momenta.unbind()), masses.unbind()
    last = len(p) - 1
    quarter, eighth = 0.25 * dt, 0.125 * dt

    def driving(index, kin):
        if index == 0:
            return 2.0 * kin - g * kT
        return p[index - 1].square() / mass[index - 1] - kT

    damping = [None] * last
    for index in range(last, -1, -1):
        if index == last:
            p[index] = p[index] + quarter * driving(index, kinetic)
        else:
            kick = quarter * driving(index, kinetic)
            p[index] = damping[index] * (damping[index] * p[index] + kick)
        if index > 0:
            damping[index - 1] = torch.exp(-eighth * p[index] / mass[index])
    factor = torch.exp(-0.5 * dt * p[0] / mass[0])
    kinetic = kinetic * factor.square()
    positions = positions + 0.5 * dt * torch.stack(p) / masses
    for index in range(last + 1):
        if index == last:
            p[index] = p[index] + quarter * driving(index, kinetic)
        else:
            kick = quarter * driving(index, kinetic)
            p[index] = damping[index] * (damping[index] * p[index] + kick)
    return factor, positions, torch.stack(p)


def heat_bath_energy(positions, momenta, masses, g, kT):
    """``_chain.py:94-105``."""
    return (momenta.square() / (2.0 * masses)).sum() + g * kT * positions[0] + kT * positions[1:].sum()


class LJModel:
    """The force evaluation of one LJ model in fp32/fp64 eager torch (reference autograd path)."""

    def __init__(self, atom_types, sigma, epsilon, cutoff, dtype):
        self.atom_types = torch.as_tensor(atom_types, dtype=torch.int64)
        self.sigma = torch.as_tensor(sigma, dtype=dtype)
        self.epsilon = torch.as_tensor(epsilon, dtype=dtype)
        self.cutoff = float(cutoff)

    def energy(self, positions, strain, cell, index):
        return op.t_lj_energy(
            positions, strain, cell, index, self.atom_types, self.sigma, self.epsilon, self.cutoff
        )


class TableModel:
    """The force evaluation of one tabulated model in fp32/fp64 eager torch."""

    def __init__(self, atom_types, table, start, cutoff, dtype):
        self.atom_types = torch.as_tensor(atom_types, dtype=torch.int64)
        self.table = torch.as_tensor(table, dtype=dtype)
        self.start = torch.as_tensor(start, dtype=dtype)
        self.cutoff = float(cutoff)
        # the spacing as the reference computes it, in the table's dtype
        self.spacing = (self.cutoff - self.start) / (self.table.shape[1] - 1)

    def energy(self, positions, strain, cell, index):
        return op.t_table_energy(positions, strain, cell, index, self.atom_types, self.table,
                                 self.start, self.spacing, self.cutoff)


class SumModel:
    """A sum of models (``potential/_sum/_potential.py:86-99``): the parts' energies added in
    part order, each part masking pairs beyond its own cutoff; the cutoff is the largest."""

    def __init__(self, *parts):
        self.parts = parts
        self.cutoff = max(part.cutoff for part in parts)

    def energy(self, positions, strain, cell, index):
        total = None
        for part in self.parts:
            e = part.energy(positions, strain, cell, index)
            total = e if total is None else total + e
        return total


class CPUList:
    """A neighbor list on the C restatement: update follows the reference's reuse rule."""

    def __init__(self, cutoff, skin, capacity_multiplier=1.25):
        self.cutoff, self.radius = float(cutoff), float(cutoff + skin)
        self.multiplier = capacity_multiplier

    def allocate(self, x_wrapped: np.ndarray, cell: np.ndarray):
        self.dims = on.grid_dims(cell.tolist(), self.radius, x_wrapped.dtype)
        self.all_pairs = on.all_pairs_path(self.dims)
        b = on.build(x_wrapped, cell, self.radius, capacity_multiplier=self.multiplier)
        self.capacity = b.capacity
        self.index = b.pairs
        self.total = b.pair_total
        self.x0, self.cell0 = x_wrapped.copy(), cell.copy()
        self.limits = np.array(self.dims, dtype=np.int32)
        self.rebuilds = 1

    def update(self, x_wrapped: np.ndarray, cell: np.ndarray) -> bool:
        rebuild, status, dims, widths = on.update_decision(
            x_wrapped, cell, self.x0, self.cell0, self.cutoff, self.radius, self.limits
        )
        assert status == 0, status
        if rebuild:
            b = on.build(
                x_wrapped, cell, self.radius, dims=dims, widths=widths,
                all_pairs=self.all_pairs, capacity=self.capacity,
            )
            assert b.status == 0
            self.index, self.total = b.pairs, b.pair_total
            self.x0, self.cell0 = x_wrapped.copy(), cell.copy()
            self.rebuilds += 1
        return bool(rebuild)


class Integrator:
    """NVE or NVT steps in eager torch on the CPU, following ``Simulation`` step order."""

    def __init__(self, model: LJModel, x, v, masses, cell, *, dt, skin, temperature=None, tau=None):
        dtype = x.dtype
        self.model, self.dtype = model, dtype
        self.cell = cell.clone()
        self.masses = masses.clone()
        self.inverse_masses = 1.0 / masses[:, None]
        self.dt = torch.tensor(dt, dtype=dtype)
        self.half_dt = torch.tensor(0.5 * dt, dtype=dtype)
        self.x, crossings = wrap(x, cell)
        self.images = crossings
        self.v = v - (masses[:, None] * v).sum(dim=0) / masses.sum()
        self.list = CPUList(model.cutoff, skin)
        self.list.allocate(self.x.numpy(), self.cell.numpy())
        self.energy, self.forces = self.evaluate()
        self.nvt = temperature is not None
        self.g = 3 * x.shape[0] - 3
        if self.nvt:
            self.kT = BOLTZMANN * float(temperature)
            self.Q = chain_masses(self.g, self.kT, float(tau))
            self.chain_x = torch.zeros(3, dtype=torch.float64)
            self.chain_p = torch.zeros(3, dtype=torch.float64)
            self.kinetic = self.kinetic_energy().double()
            self.dt64 = torch.tensor(dt, dtype=torch.float64)

    def evaluate(self):
        strain = self.x.new_zeros((3, 3))
        g = torch.func.grad_and_value(self.model.energy, argnums=0)
        gx, e = g(self.x, strain, self.cell, torch.as_tensor(self.list.index).long())
        return e, -gx

    def kinetic_energy(self):
        return 0.5 * (self.masses[:, None] * self.v.square()).sum()

    def _chain(self):
        factor, self.chain_x, self.chain_p = chain_half_step(
            self.kinetic, self.chain_x, self.chain_p, self.Q, self.g, self.kT, self.dt64
        )
        self.v = self.v * factor.to(self.dtype)
        self.kinetic = self.kinetic * factor.square()

    def step(self):
        if self.nvt:
            self._chain()
        self.v = self.v + self.half_dt * self.forces * self.inverse_masses
        self.x, crossings = wrap(self.x + self.dt * self.v, self.cell)
        self.images = self.images + crossings
        self.list.update(self.x.numpy(), self.cell.numpy())
        self.energy, self.forces = self.evaluate()
        self.v = self.v + self.half_dt * self.forces * self.inverse_masses
        if self.nvt:
            self.kinetic = self.kinetic_energy().double()
            self._chain()

    def total_energy(self):
        return float(self.energy) + float(self.kinetic_energy())

    def conserved_energy(self):
        bath = heat_bath_energy(self.chain_x, self.chain_p, self.Q, self.g, self.kT)
        return float(self.energy.double() + self.kinetic_energy().double() + bath)
