"""The state of a simulation (``md/_state.py``): one frame, detached from the simulation."""

from typing import NamedTuple

from torch import Tensor


class ChainState(NamedTuple):
    """A Nose-Hoover chain: positions [3], dimensionless, and momenta [3] in kJ ps/mol, fp64."""

    positions: Tensor
    momenta: Tensor


class State(NamedTuple):
    """A copy of a simulation's frame: positions [N, 3] in [0, L), velocities, forces, cell,
    images [N, 3] int64 (unwrapped = positions + images * cell), step, and the thermostat of an
    NVT run. Pass these back to ``nve``/``nvt`` to continue a run."""

    positions: Tensor
    velocities: Tensor
    forces: Tensor
    cell: Tensor
    images: Tensor
    step: int
    thermostat: ChainState | None = None
    barostat: None = None


__all__ = ("ChainState", "State")
