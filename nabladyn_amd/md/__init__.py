"""Molecular dynamics on the GPU: each step replays one captured HIP graph (``md/__init__.py``)."""

from ._simulation import FIELDS, Simulation, nve, nvt
from ._state import ChainState, State
from ._velocities import maxwell_boltzmann

__all__ = ("nve", "nvt", "maxwell_boltzmann", "Simulation", "FIELDS", "State", "ChainState")
