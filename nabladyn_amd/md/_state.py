"""The state of a simulation (``md/_state.py``): one frame, detached from the simulation."""

from typing import NamedTuple

from torch import Tensor


class ChainState(NamedTuple):
    """A Nose-Hoover chain: positions [3], dimensionless, and momenta [3] in kJ ps/mol, fp64."""

    positions: Tensor
    momenta: Tensor


class BarostatState(NamedTuple):
    """An isotropic barostat: the box momentum [] in kJ ps/mol, float64, and its chain."""

    momentum: Tensor
    chain: ChainState


class State(NamedTuple):
    """A copy of a simulation's frame: positions [N, 3] in [0, L), velocities, forces, cell,
    images [N, 3] int64 (unwrapped = positions + images * cell), step, the thermostat of an NVT
    or NPT run and the barostat of an NPT run. Pass these back to ``nve``/``nvt``/``npt`` to
    continue a run."""

    positions: Tensor
    velocities: Tensor
    forces: Tensor
    cell: Tensor
    images: Tensor
    step: int
    thermostat: ChainState | None = None
    barostat: BarostatState | None = None


__all__ = ("BarostatState", "ChainState", "State")
