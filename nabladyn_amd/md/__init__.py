"""Molecular dynamics on the GPU: each step replays one captured HIP graph (``md/__init__.py``)."""

from ._batch import SimulationBatch, batch_key
from ._simulation import FIELDS, Simulation, npt, nve, nvt
from ._state import BarostatState, ChainState, State
from ._velocities import maxwell_boltzmann

__all__ = ("nve", "nvt", "npt", "maxwell_boltzmann", "Simulation", "SimulationBatch", "batch_key",
           "FIELDS", "State", "ChainState", "BarostatState")
