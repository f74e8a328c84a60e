"""Several simulations stepped together: one launch per kernel of the step for the whole batch.

A member computes the same bits in a batch as alone (``Simulation.run``): each kernel gives every
member the blocks its own step launches, and each block runs the single-system code on its
member's data. The force kernel, the only one that depends on the potential, is launched once for
each kind the members bind (LJ or a table, of one type or several). What a batch saves is
launches: a step of ``k`` small systems costs the step's 11 (or so) launches once instead of
``k`` times, so many small systems run at the rate of one large one.

There is no reference counterpart: the reference compiles and replays one system's step
(``md/_simulation.py:316-364``); a batch is that step for every member at once.
"""

from collections.abc import Sequence
from numbers import Integral

import torch

from .. import _native
from ._simulation import ListState, MDList, MDSettings, MDSystem, Simulation, _count

_KERNELS = 13
# the launches of a step (ndn_md_batch.grids): the forces, one per kind of potential, come eighth
# to eleventh: one-type LJ, typed LJ, typed table, one-type table; the finish is last
FORCE_LAUNCHES = {"lj": 8, "typed_lj": 9, "typed_table": 10, "table": 11}
_MAX_MEMBERS = 256


class _Member(_native.ctypes.Structure):
    """``ndn_md_member``."""

    _fields_ = [
        ("system", _native.ctypes.POINTER(MDSystem)),
        ("settings", _native.ctypes.POINTER(MDSettings)),
        ("potential", _native.ctypes.POINTER(_native.Potential)),
        ("list", _native.ctypes.POINTER(ListState)),
        ("md_list", _native.ctypes.POINTER(MDList)),
    ]


class _Batch(_native.ctypes.Structure):
    """``ndn_md_batch``."""

    _fields_ = [
        ("table", _native.ctypes.c_void_p), ("table_bytes", _native.ctypes.c_int64),
        ("dtype", _native.ctypes.c_int32), ("count", _native.ctypes.c_int32),
        ("kinds", _native.ctypes.c_int32), ("group", _native.ctypes.c_int32),
        ("virial", _native.ctypes.c_int32), ("continues", _native.ctypes.c_int32),
        ("members_offset", _native.ctypes.c_int64),
        ("potentials_offset", _native.ctypes.c_int64), ("blocks_offset", _native.ctypes.c_int64),
        ("offsets", _native.ctypes.c_int64 * _KERNELS), ("grids", _native.ctypes.c_int64 * _KERNELS),
        ("lds", _native.ctypes.c_int64 * _KERNELS),
    ]


def batch_key(simulation: Simulation) -> tuple:
    """Simulations with equal keys can share a batch: the same device and dtype, the same force
    schedule and virial setting (the potentials may differ: each kind present gets its own force
    launch)."""
    return (str(simulation._device), simulation._dtype, simulation.group,
            bool(simulation._virial_every_step))


class SimulationBatch:
    """Steps ``simulations`` (1 to 256 with one ``batch_key``) together on the current stream.

    Each member keeps its own state, step count and status; ``advance`` moves them all. The batch
    reads the members' buffers as they are now: make a new batch after a member is rebuilt with
    other capacities (a simulation keeps its sizes for its life, so this never happens within
    ``advance``). Raises ``ValueError`` for an empty, oversized or mixed set."""

    def __init__(self, simulations: Sequence[Simulation]):
        simulations = list(simulations)
        if not simulations or len(simulations) > _MAX_MEMBERS:
            raise ValueError(f"a batch holds 1 to {_MAX_MEMBERS} simulations")
        if len({id(s) for s in simulations}) != len(simulations):
            raise ValueError("a simulation appears twice in the batch")
        keys = {batch_key(s) for s in simulations}
        if len(keys) != 1:
            raise ValueError(
                "the simulations of a batch must share the device, dtype, force group and "
                f"virial setting; got {sorted(map(str, keys))}")
        for simulation in simulations:
            simulation._require()
            if simulation._pot.next:
                raise ValueError("a batch steps simulations of one LennardJones or Tabulated "
                                 "model each; a sum of two is stepped alone")
        first = simulations[0]
        self.simulations = simulations
        self._device = first._device
        members = (_Member * len(simulations))()
        for member, s in zip(members, simulations):
            member.system = _native.ctypes.pointer(s._system)
            member.settings = _native.ctypes.pointer(s._settings)
            member.potential = _native.ctypes.pointer(s._pot)
            member.list = _native.ctypes.pointer(s._neighbors.state)
            member.md_list = _native.ctypes.pointer(s._neighbors.native)
        self._members = members
        batch = _Batch()
        batch.dtype = _native.dtype_tag(first._dtype)
        stream = _native.stream(self._device)
        _native.call("md_batch_prepare", _native.ctypes.byref(batch), members, len(simulations),
                     stream)
        self.table = torch.empty(batch.table_bytes, dtype=torch.uint8, device=self._device)
        batch.table = self.table.data_ptr()
        _native.call("md_batch_prepare", _native.ctypes.byref(batch), members, len(simulations),
                     stream)
        self._batch = batch

    def __len__(self) -> int:
        return len(self.simulations)

    @property
    def grids(self) -> tuple[int, ...]:
        """Blocks of each of the step's launches over the whole batch."""
        return tuple(int(v) for v in self._batch.grids)

    def advance(self, steps: Integral) -> None:
        """Queues ``steps`` batched steps on the current stream and returns at once (no sync and
        no status read; each member's ``check`` reads its status)."""
        steps = _count("steps", steps)
        for simulation in self.simulations:
            simulation._require()
        stream = _native.stream(self._device)
        handle = _native.ctypes.byref(self._batch)
        step = _native.lib().ndn_md_batch_step
        # the first step of a call runs the prologue; each later one was opened by the step
        # before it (its finish)
        for k in range(steps):
            self._batch.continues = 1 if k > 0 else 0
            _native.check("md_batch_step", step(handle, stream))
        self._batch.continues = 0
        for simulation in self.simulations:
            simulation._step += steps

    def check(self) -> None:
        """Reads every member's status (syncs); raises ValueError for the first that failed."""
        for simulation in self.simulations:
            simulation.check()


__all__ = ("SimulationBatch", "batch_key")
