"""The simulation (``md/_simulation.py``): a resident frame, its neighbor list, and one HIP graph
holding a whole MD step, replayed by ``run``.

Where the reference compiles the ensemble's halves with Inductor and captures them around the
neighbor operator, here the step is the library's ``ndn_md_step`` — prologue, first half with
the reuse check, gated rebuild, fused force and second kick, finish — captured once by
``torch.cuda.CUDAGraph`` (a hipGraph on ROCm). The neighbor status is read every 1,000 steps,
as in the reference.
"""

import math
import sys
from numbers import Integral, Real

import torch
from torch import Tensor

from .. import _native, space
from ..potential import Potential
from ..units import BOLTZMANN
from ._state import BarostatState, ChainState, State

_STATUS_INTERVAL = 1000
# Steps captured in the multi-step graph: one replay launches this many steps, so the host
# issues one call per 25 steps (small systems' steps take less GPU time than a host replay).
_CHUNK_STEPS = 25
# Skin of the pruned list the force kernel reads (nm): the pairs of the full list within
# cutoff + INNER_SKIN, refreshed whenever an atom has moved half of it. Only the force kernel's
# work depends on it (the full list, its skin and its rebuild rule are the reference's). At
# 2,097,152 atoms of argon (cutoff 1.0 nm, skin 0.2 nm) the step took 0.485 / 0.468 / 0.468 /
# 0.467 / 0.455 / 0.467 / 0.491 ms at 0.06 / 0.07 / 0.08 / 0.09 / 0.10 / 0.12 / 0.14 nm.
INNER_SKIN = 0.10
# At most this many atoms, a simulation keeps no pruned list unless asked for one.
_SMALL_SYSTEM = 2048
FIELDS = ("potential_energy", "kinetic_energy", "total_energy", "temperature", "pressure")


class MDSystem(_native.ctypes.Structure):
    """``ndn_md_system``."""

    _fields_ = [
        ("num_atoms", _native.ctypes.c_int64), ("positions", _native.ctypes.c_void_p),
        ("velocities", _native.ctypes.c_void_p), ("forces", _native.ctypes.c_void_p),
        ("images", _native.ctypes.c_void_p), ("masses", _native.ctypes.c_void_p),
        ("inverse_masses", _native.ctypes.c_void_p), ("cell", _native.ctypes.c_void_p),
        ("energy", _native.ctypes.c_void_p), ("virial", _native.ctypes.c_void_p),
        ("thermostat", _native.ctypes.c_void_p), ("chain_masses", _native.ctypes.c_void_p),
        ("kinetic", _native.ctypes.c_void_p), ("partials", _native.ctypes.c_void_p),
        ("partial_slots", _native.ctypes.c_int64), ("coefficients", _native.ctypes.c_void_p),
        ("coefficient_bytes", _native.ctypes.c_int64), ("barostat", _native.ctypes.c_void_p),
    ]


class MDSettings(_native.ctypes.Structure):
    """``ndn_md_settings``."""

    _fields_ = [
        ("timestep", _native.ctypes.c_double), ("kT", _native.ctypes.c_double),
        ("degrees_of_freedom", _native.ctypes.c_int64), ("thermostat", _native.ctypes.c_int),
        ("virial", _native.ctypes.c_int), ("pressure", _native.ctypes.c_double),
        ("box_mass", _native.ctypes.c_double), ("alpha", _native.ctypes.c_double),
        ("continues", _native.ctypes.c_int),
    ]


class ListState(_native.ctypes.Structure):
    """``ndn_list_state``."""

    _fields_ = [
        ("cutoff", _native.ctypes.c_double), ("radius", _native.ctypes.c_double),
        ("reference_positions", _native.ctypes.c_void_p), ("reference_cell", _native.ctypes.c_void_p),
        ("limits", _native.ctypes.c_void_p), ("reuse", _native.ctypes.c_void_p),
        ("rebuild", _native.ctypes.c_void_p), ("rebuild_count", _native.ctypes.c_void_p),
    ]


class MDList(_native.ctypes.Structure):
    """``ndn_md_list``."""

    _fields_ = [
        ("dims", _native.ctypes.c_int32 * 3), ("group", _native.ctypes.c_int32),
        ("widths", _native.ctypes.c_double * 3), ("stride", _native.ctypes.c_int64),
        ("bin_counts", _native.ctypes.c_void_p), ("bin_offsets", _native.ctypes.c_void_p),
        ("bin_of", _native.ctypes.c_void_p), ("unordered", _native.ctypes.c_void_p),
        ("order", _native.ctypes.c_void_p), ("counts", _native.ctypes.c_void_p),
        ("index", _native.ctypes.c_void_p), ("ids", _native.ctypes.c_void_p),
        ("types", _native.ctypes.c_void_p), ("positions4", _native.ctypes.c_void_p),
        ("scratch", _native.ctypes.c_void_p), ("scan_temp", _native.ctypes.c_void_p),
        ("scan_temp_bytes", _native.ctypes.c_int64), ("status", _native.ctypes.c_void_p),
        ("prepare_dims", _native.ctypes.c_void_p), ("prepare_widths", _native.ctypes.c_void_p),
        ("prepare_total", _native.ctypes.c_void_p),
        ("inner_radius", _native.ctypes.c_double), ("inner_counts", _native.ctypes.c_void_p),
        ("inner_index", _native.ctypes.c_void_p), ("prune_positions", _native.ctypes.c_void_p),
        ("brick", _native.ctypes.c_int32 * 3), ("bricks", _native.ctypes.c_int64),
        ("halo_capacity", _native.ctypes.c_int64), ("halo_slots", _native.ctypes.c_void_p),
        ("halo_runs", _native.ctypes.c_void_p), ("halo_counts", _native.ctypes.c_void_p),
        ("grid_widths", _native.ctypes.c_void_p),
    ]


_GROUPS = (4, 8, 16)
_INVALID_CELL, _NON_FINITE, _OVERFLOW, _FULL_OVERFLOW, _HALO_OVERFLOW = 1, 2, 8, 16, 32
_CELL_SHRANK = 64
# NPT: the sub-bins are chosen wider than half the list's radius by this factor, so the cell may
# shrink by as much (about 15% in volume) before the fixed grid stops covering the radius
NPT_GRID_MARGIN = 0.05
# barostat slots (fp64): box momentum, its chain's q [3], p [3] and masses [3], the master cell
_BOX_P, _BOX_Q, _BOX_PC, _BOX_MASS, _MASTER_CELL = 0, 1, 4, 7, 10
_HALO_RUNS = 128  # halo run slots per brick
# brick shapes (sub-bins along x, y, z), largest first: the first whose expected halo fits wins
_BRICKS = ((4, 4, 8), (4, 2, 8), (2, 2, 8), (2, 2, 4), (2, 1, 4), (1, 1, 4), (1, 1, 2),
           (1, 1, 1))


def brick_shapes(dims: list[int], num_atoms: int, dtype: torch.dtype, reserve: int = 0):
    """The brick shapes worth trying for this grid, largest first: those whose expected halo
    (atoms per sub-bin times its sub-bins, with a tenth to spare) fits the LDS budget less
    `reserve` positions' worth kept for the potential."""
    per_bin = num_atoms / math.prod(dims)
    limit = halo_limit(dtype) - reserve
    shapes = []
    for shape in _BRICKS:
        halo = math.prod(min(b + 4, d) for b, d in zip(shape, dims))
        if 1.1 * halo * per_bin <= limit:
            shapes.append(shape)
    return shapes or [_BRICKS[-1]]


def potential_reserve(model) -> int:
    """Halo positions' worth of LDS the force kernel stages for the model beside the halo: a
    one-type table's segments, two 4-vectors each (the LDS budget of two blocks per CU holds
    both)."""
    from ..potential import Tabulated

    if isinstance(model, Tabulated) and model.table.shape[0] == 1:
        return 2 * (int(model.table.shape[1]) - 1)
    return 0


def halo_limit(dtype: torch.dtype) -> int:
    """The largest halo the force kernel's LDS holds on the current device."""
    query = _native.lib().ndn_md_halo_limit
    query.restype = _native.ctypes.c_int64
    return int(query(_native.dtype_tag(dtype)))


def sub_bin_dims(lengths: list[float], radius: float, dtype: torch.dtype,
                 margin: float = 0.0) -> list[int]:
    """Sub-bins per axis, each at least radius / 2 wide with a margin for the rounding of binning
    (so every pair within the radius lies within two sub-bins along each axis); ``margin``
    widens them further by that fraction (a cell that may shrink)."""
    eps = torch.finfo(dtype).eps
    half = 0.5 * radius * (1.0 + 1e-6) * (1.0 + margin)
    return [max(1, math.floor(length / (half + 4 * eps * length))) for length in lengths]


class _MDNeighbors:
    """The MD path's neighbor state: the reuse rule's tensors and the full list in slot order."""

    def __init__(self, cutoff, skin, multiplier, positions, cell, types, group_hint=None,
                 inner_skin=None, reserve=0, grid_margin=0.0, halo_margin=None):
        from ..neighbors._neighbor_list import grid_dims

        # LDS kept for the potential beside the halo, in halo positions
        self.reserve = int(reserve)
        self.cutoff, self.skin, self.multiplier = float(cutoff), float(skin), float(multiplier)
        # the pruned list, when its skin is positive and below the list's own
        self.inner_skin = (float(inner_skin) if inner_skin is not None and 0 < inner_skin < skin
                           else None)
        n, device, dtype = positions.shape[0], positions.device, positions.dtype
        lengths = [float(v) for v in cell.tolist()]
        ints = {"dtype": torch.int32, "device": device}
        self.limits = torch.tensor(grid_dims(lengths, self.radius, dtype), **ints)
        dims = sub_bin_dims(lengths, self.radius, dtype, grid_margin)
        # headroom of the halos over the first build's largest (half the rows' by default)
        self.halo_margin = (1.0 + 0.5 * (self.multiplier - 1.0) if halo_margin is None
                            else float(halo_margin))
        bins = math.prod(dims)
        if bins >= 2**31 - 1:
            raise ValueError("the grid must have fewer than 2**31 - 1 bins")
        self.reference_positions = positions.clone()
        self.reference_cell = torch.full((3,), math.inf, dtype=dtype, device=device)
        self.parameters = torch.zeros(4, dtype=dtype, device=device)
        self.rebuild = torch.zeros(1, **ints)
        self.rebuild_count = torch.zeros(1, **ints)
        scan = _native.ctypes.c_int64()
        _native.call("scan_temp_bytes", _native.ctypes.c_int64(bins + 1), _native.ctypes.byref(scan))
        scratch = _native.lib().ndn_md_scratch_bytes
        scratch.restype = _native.ctypes.c_int64
        self.buffers = {
            "bin_counts": torch.zeros(bins + 1, **ints),
            "bin_offsets": torch.zeros(bins + 1, **ints),
            "bin_of": torch.zeros(n, **ints),
            "unordered": torch.zeros(n, **ints),
            "order": torch.zeros(n, **ints),
            "counts": torch.zeros(n, **ints),
            "index": torch.zeros(1, dtype=torch.int16, device=device),
            "ids": torch.arange(n, **ints),
            "types": types.clone() if types is not None else None,
            # x, y, z, type of each slot (row n spare)
            "positions4": torch.zeros((n + 1, 4), dtype=dtype, device=device),
            "scratch": torch.zeros(int(scratch(_native.dtype_tag(dtype), _native.ctypes.c_int64(n))),
                                   dtype=torch.uint8, device=device),
            "scan_temp": torch.zeros(max(scan.value, 4096), dtype=torch.uint8, device=device),
            "status": torch.zeros(2, **ints),
            "prepare_dims": torch.zeros(3, **ints),
            "prepare_widths": torch.zeros(3, dtype=dtype, device=device),
            "prepare_total": torch.zeros(1, **ints),
            # the sub-bin widths (cell / dims) the kernels read; an NPT step rewrites them
            "grid_widths": torch.tensor([L / d for L, d in zip(lengths, dims)], dtype=dtype,
                                        device=device),
        }
        if self.inner_skin is not None:
            b_ = self.buffers
            b_["inner_counts"] = torch.zeros(n, **ints)
            b_["inner_index"] = torch.zeros(1, dtype=torch.int16, device=device)
            b_["prune_positions"] = torch.zeros((n, 3), dtype=dtype, device=device)
        b = self.buffers
        self.native = MDList(
            dims=(_native.ctypes.c_int32 * 3)(*dims),
            group=8 if group_hint is None else group_hint,
            widths=(_native.ctypes.c_double * 3)(*[L / d for L, d in zip(lengths, dims)]),
            stride=0,
            **{k: _native.ptr(v) for k, v in b.items() if k not in ("scan_temp", "inner_index")},
            scan_temp=_native.ptr(b["scan_temp"]),
            scan_temp_bytes=b["scan_temp"].numel(),
            inner_radius=self.cutoff + self.inner_skin if self.inner_skin is not None else 0.0,
        )
        self.state = ListState(
            cutoff=self.cutoff, radius=self.radius,
            reference_positions=_native.ptr(self.reference_positions),
            reference_cell=_native.ptr(self.reference_cell), limits=_native.ptr(self.limits),
            reuse=_native.ptr(self.parameters), rebuild=_native.ptr(self.rebuild),
            rebuild_count=_native.ptr(self.rebuild_count),
        )
        self.dims = dims
        self.dtype = dtype
        self.set_bricks(brick_shapes(dims, n, dtype, self.reserve)[0])

    @property
    def radius(self) -> float:
        return self.cutoff + self.skin

    def set_bricks(self, shape) -> None:
        """Bricks of `shape` sub-bins; their halos unsized until `size_halo`."""
        device = self.buffers["counts"].device
        bricks = math.prod(-(-d // b) for d, b in zip(self.dims, shape))
        self.brick = tuple(shape)
        self.bricks = bricks
        self.buffers["halo_runs"] = torch.zeros(bricks * _HALO_RUNS, dtype=torch.int32, device=device)
        self.buffers["halo_counts"] = torch.zeros(bricks, dtype=torch.int32, device=device)
        self.buffers["halo_slots"] = torch.zeros(1, dtype=torch.int32, device=device)
        self.native.brick = (_native.ctypes.c_int32 * 3)(*shape)
        self.native.bricks = bricks
        self.native.halo_capacity = 0
        for name in ("halo_runs", "halo_counts", "halo_slots"):
            setattr(self.native, name, _native.ptr(self.buffers[name]))

    @property
    def halo_needed(self) -> int:
        """The largest halo of the last build."""
        return int(self.buffers["scan_temp"][656:660].view(torch.int32).item())

    def size_halo(self, needed: int) -> None:
        """Room for halos of `needed` atoms with a margin (half the rows' multiplier's), up to
        what LDS holds."""
        margin = self.halo_margin
        limit = halo_limit(self.dtype)
        # within the budget less the potential's share when the halo fits it (else as LDS allows)
        room = limit - self.reserve if needed <= limit - self.reserve else limit
        capacity = min(max(1, math.ceil(margin * needed)), room)
        self.buffers["halo_slots"] = torch.zeros(self.bricks * capacity, dtype=torch.int32,
                                                 device=self.buffers["counts"].device)
        self.native.halo_slots = _native.ptr(self.buffers["halo_slots"])
        self.native.halo_capacity = capacity

    @property
    def halo_capacity(self) -> int:
        return int(self.native.halo_capacity)

    def size_rows(self, max_count: int) -> None:
        """Allocate the rows for `max_count` neighbors times the capacity multiplier."""
        # a multiple of 64: a row's last block of 4G entries (G <= 16 lanes per row) stays in it
        stride = max(64, math.ceil(self.multiplier * max_count))
        stride = (stride + 63) // 64 * 64
        n = self.buffers["counts"].numel()
        # 16-bit local indices into the brick's halo
        self.buffers["index"] = torch.zeros(n * stride, dtype=torch.int16,
                                            device=self.buffers["counts"].device)
        if self.inner_skin is not None:
            # a pruned row is a subset of its full row: the same stride always holds it
            self.buffers["inner_index"] = torch.zeros_like(self.buffers["index"])
            self.native.inner_index = _native.ptr(self.buffers["inner_index"])
        self.native.index = _native.ptr(self.buffers["index"])
        self.native.stride = stride

    @property
    def stride(self) -> int:
        return int(self.native.stride)

    @property
    def _rebuild_count(self) -> int:
        return int(self.rebuild_count.item())

    @property
    def num_pairs(self) -> int:
        """Pairs within the list radius at the last build (each counted once)."""
        return int(self.buffers["counts"].sum(dtype=torch.int64).item()) // 2

    def _max_moved(self) -> float:
        """The largest squared displacement measured at the reuse checks since the last prologue
        (a step's own, or the first step of a captured run), in nm^2."""
        raw = self.buffers["scan_temp"][64:64 + 8 * 64].view(torch.int64).max().item()
        import struct

        if self.reference_positions.dtype == torch.float32:
            return struct.unpack("<f", struct.pack("<I", raw & 0xFFFFFFFF))[0]
        return struct.unpack("<d", struct.pack("<q", raw))[0]

    def _check_status(self) -> None:
        flags, needed = self.buffers["status"].tolist()
        if flags:
            causes = []
            if flags & _INVALID_CELL:
                causes.append("every cell length must be finite and above twice the cutoff")
            if flags & _NON_FINITE:
                causes.append("positions must be finite")
            if flags & _FULL_OVERFLOW:
                causes.append(
                    f"a neighbor row needs {needed} entries but holds {self.stride}; "
                    "raise capacity_multiplier",
                )
            if flags & _HALO_OVERFLOW:
                causes.append(
                    f"a brick's halo needs {self.halo_needed} atoms but holds "
                    f"{self.halo_capacity}; raise capacity_multiplier",
                )
            if flags & _CELL_SHRANK:
                causes.append(
                    "the cell shrank below what the simulation's sub-bin grid covers (more "
                    f"than about {NPT_GRID_MARGIN:.0%} in length); continue from state() in a "
                    "new simulation",
                )
            raise ValueError("; ".join(causes))

    @property
    def refresh_count(self) -> int:
        """Refreshes of the pruned list so far (builds included)."""
        return int(self.buffers["scan_temp"][652:656].view(torch.int32).item())

    def atom_pairs(self, pruned: bool = False):
        """The full list (or the pruned list) as [P, 2] atom-id pairs (both orientations), on
        the host; for tests."""
        b = self.buffers
        counts = b["inner_counts" if pruned else "counts"].cpu()
        stride = self.stride
        slots = torch.full((counts.numel() * stride,), -1, dtype=torch.int32,
                           device=b["counts"].device)
        _native.call("md_list_entries", _native.ctypes.byref(self.native),
                     _native.ctypes.c_int64(counts.numel()), 1 if pruned else 0,
                     _native.ptr(slots), _native.stream(slots.device))
        # a row's slots, then -1 past its neighbors (a pruned row's count includes the
        # sentinels that even out its lanes' shares)
        index = slots.view(-1, stride).cpu().long()
        ids = b["ids"].cpu().long()
        rows, columns = torch.nonzero(index >= 0, as_tuple=True)
        return torch.stack([ids[rows], ids[index[rows, columns]]], dim=1)


def _positive(name: str, value) -> float:
    if isinstance(value, bool) or not isinstance(value, Real) or not math.isfinite(value) or value <= 0:
        raise ValueError(f"{name} must be a finite positive number")
    return float(value)


def _count(name: str, value) -> int:
    if isinstance(value, bool) or not isinstance(value, Integral) or value < 0:
        raise ValueError(f"{name} must be a nonnegative integer")
    return int(value)


def _check_frame(model, positions, velocities, masses, cell) -> None:
    """``md/_simulation.py:66-88``."""
    num_atoms, dtype, device = model.num_atoms, model.dtype, model.device
    for name, value, shape in (
        ("positions", positions, (num_atoms, 3)), ("velocities", velocities, (num_atoms, 3)),
        ("masses", masses, (num_atoms,)), ("cell", cell, (3,)),
    ):
        if not isinstance(value, Tensor) or value.shape != shape or value.dtype != dtype or value.device != device:
            raise ValueError(
                f"{name} must be a {list(shape)} tensor of {dtype} on {device} to match the model",
            )


def _check_npt_settings(temperature, pressure, thermostat_timescale, barostat_timescale) -> None:
    """``md/_ensembles/_npt.py:40-56``: temperature and timescales finite and positive, the
    pressure finite."""
    for name, value in (("temperature", temperature), ("thermostat_timescale", thermostat_timescale),
                        ("barostat_timescale", barostat_timescale)):
        _positive(name, value)
    if isinstance(pressure, bool) or not isinstance(pressure, Real) or not math.isfinite(pressure):
        raise ValueError("pressure must be a finite number")


def _check_barostat(barostat, device) -> None:
    momentum = barostat.momentum if isinstance(barostat, BarostatState) else None
    if (not isinstance(momentum, Tensor) or momentum.shape != () or momentum.dtype != torch.float64
            or momentum.device != device):
        raise ValueError(f"barostat.momentum must be a [] float64 tensor on {device}")
    for field, value in zip(("positions", "momenta"), barostat.chain):
        if (not isinstance(value, Tensor) or value.shape != (3,) or value.dtype != torch.float64
                or value.device != device):
            raise ValueError(f"barostat.chain.{field} must be a [3] float64 tensor on {device}")


def _check_chain(chain, device) -> None:
    for field, value in zip(("positions", "momenta"), chain):
        if not isinstance(value, Tensor) or value.shape != (3,) or value.dtype != torch.float64 or value.device != device:
            raise ValueError(f"thermostat.{field} must be a [3] float64 tensor on {device}")


class Simulation:
    """Molecular dynamics of a potential on the GPU; build it with ``nve`` or ``nvt``.

    It owns its frame, a neighbor list at the model's cutoff (with a full list for the fused
    force kernel) and HIP graphs of one step and of 25, which ``run`` replays. With ``graphs``
    False it captures none and launches each step as it comes (for a simulation stepped in a
    ``SimulationBatch``, which saves the captures). ``inner_skin`` (nm) is the skin of the pruned
    list the force kernel reads: "auto" takes ``INNER_SKIN`` above 2,048 atoms when the list's
    skin is larger than it, and none otherwise (the force kernel then reads the full rows); None
    keeps none; a number must lie strictly between 0 and the list's skin, or none is kept.
    """

    def __init__(self, model: Potential, positions: Tensor, velocities: Tensor, masses: Tensor,
                 cell: Tensor, *, timestep: Real, skin: Real, capacity_multiplier: Real,
                 images: Tensor | None, step: Integral, temperature: Real | None = None,
                 thermostat_timescale: Real | None = None, thermostat: ChainState | None = None,
                 virial_every_step: bool = False, force_group: int | None = None,
                 inner_skin: Real | None | str = "auto", graphs: bool = True,
                 pressure: Real | None = None, barostat_timescale: Real | None = None,
                 barostat: BarostatState | None = None):
        if not isinstance(model, Potential):
            raise TypeError("model must be a Potential")
        if not model._fused:
            raise TypeError(
                "the MD step evaluates a LennardJones or Tabulated model, or a sum of two, with "
                "the library's kernels; another energy is not supported here")
        num_atoms, device, dtype = model.num_atoms, model.device, model.dtype
        _check_frame(model, positions, velocities, masses, cell)
        if images is None:
            images = torch.zeros((num_atoms, 3), dtype=torch.int64, device=device)
        elif not isinstance(images, Tensor) or images.shape != (num_atoms, 3) or images.dtype != torch.int64 or images.device != device:
            raise ValueError(f"images must be a [{num_atoms}, 3] int64 tensor on {device}")
        timestep = _positive("timestep", timestep)
        self._step = _count("step", step)
        velocities, masses = velocities.detach(), masses.detach().clone()
        if not bool(torch.isfinite(velocities).all()):
            raise ValueError("velocities must be finite")
        if not bool((torch.isfinite(masses) & (masses > 0)).all()):
            raise ValueError("masses must be finite and positive")
        self._nvt = temperature is not None
        self._npt = pressure is not None
        if self._npt:
            _check_npt_settings(temperature, pressure, thermostat_timescale, barostat_timescale)
            if num_atoms < 2:
                raise ValueError("NPT needs at least 2 atoms, for 3N - 3 degrees of freedom")
        if self._nvt:
            _positive("temperature", temperature)
            _positive("thermostat_timescale", thermostat_timescale)
            if num_atoms < 2:
                raise ValueError("NVT needs at least 2 atoms, for 3N - 3 degrees of freedom")
        cell = cell.detach().clone().contiguous()
        wrapped, crossings = space.wrap(positions.detach().contiguous(), cell)
        if not bool(torch.isfinite(wrapped).all()):
            raise ValueError("positions must be finite")
        lengths = [float(v) for v in cell.tolist()]
        if not all(math.isfinite(v) and v > 2 * model.cutoff for v in lengths):
            raise ValueError("every cell length must be finite and above twice the cutoff")
        column = masses[:, None]
        velocities = velocities - (column * velocities).sum(dim=0) / masses.sum()
        self._model, self._timestep = model, timestep
        self._dtype, self._device = dtype, device
        self._num_atoms = num_atoms
        self._degrees_of_freedom = 3 * num_atoms - 3
        # NPT drives the box with the virial of every step
        self._virial_every_step = bool(virial_every_step) or self._npt
        wide = {"dtype": torch.float64, "device": device}
        # resident tensors in slot order: their addresses are fixed for the graph
        self.positions = wrapped.contiguous()
        self.velocities = velocities.contiguous().clone()
        self.forces = torch.zeros_like(self.positions)
        self.images = (images + crossings).contiguous()
        self.cell = cell
        self.slot_masses = masses.contiguous().clone()
        self.inverse_masses = (1.0 / masses).contiguous()
        self._masses = masses
        self.energy = torch.zeros((), dtype=dtype, device=device)
        self.virial = torch.zeros((3, 3), dtype=dtype, device=device)
        # chain q [3], p [3], the pending closing factor, the next opening factor, and the
        # chain after the next step's opening half step (q [3], p [3])
        self.thermostat = torch.zeros(14, **wide)
        self.thermostat[6] = 1.0
        self.thermostat[7] = 1.0
        self.kinetic = torch.zeros(1, **wide)
        self.settings = {"ensemble": "NPT" if self._npt else "NVT" if self._nvt else "NVE"}
        if self._nvt:
            self._kT = BOLTZMANN * float(temperature)
            tau = float(thermostat_timescale)
            base = self._kT * tau * tau
            self.chain_masses = torch.tensor([self._degrees_of_freedom * base, base, base], **wide)
            if thermostat is not None:
                _check_chain(thermostat, device)
                self.thermostat[0:3] = thermostat[0]
                self.thermostat[3:6] = thermostat[1]
            self.settings.update(temperature=float(temperature), thermostat_timescale=tau)
        else:
            self._kT = 0.0
            self.chain_masses = torch.zeros(3, **wide)
        self.barostat = None
        self._box_mass = self._alpha = self._pressure = 0.0
        if self._npt:
            # md/_ensembles/_npt.py:114-147: box mass 3 (N + 1) kT tau_b^2, the barostat's chain
            # of one degree of freedom with masses kT tau_b^2, alpha = 1 + 1/N, a master cell
            tau_b = float(barostat_timescale)
            self._pressure = float(pressure)
            self._box_mass = 3 * (num_atoms + 1) * self._kT * tau_b * tau_b
            self._alpha = 1.0 + 1.0 / num_atoms
            self.barostat = torch.zeros(16, **wide)
            self.barostat[_BOX_MASS:_BOX_MASS + 3] = self._kT * tau_b * tau_b
            self.barostat[_MASTER_CELL:_MASTER_CELL + 3] = cell.double()
            self.barostat[13] = 1.0
            self.barostat[14] = 1.0
            if barostat is not None:
                _check_barostat(barostat, device)
                self.barostat[_BOX_P] = barostat.momentum
                self.barostat[_BOX_Q:_BOX_Q + 3] = barostat.chain.positions
                self.barostat[_BOX_PC:_BOX_PC + 3] = barostat.chain.momenta
            self.settings.update(pressure=self._pressure, barostat_timescale=tau_b)
            # the pruned list's rule assumes a fixed cell between refreshes
            inner_skin = None
        pot, self._keep = model._native()
        self._pot = pot
        types = model.atom_types if getattr(model, "_num_types", 1) > 1 else None
        if inner_skin == "auto":
            # The pruned list pays for its refreshes above a few thousand atoms when the list's
            # skin leaves room for it: at the default skin of 0.1 nm a pruned skin of 0.1 nm would
            # prune nothing, and one of 0.05 nm saves too few pairs to pay for refreshes every
            # few steps, so a list skin of 0.1 nm or less keeps none. A small system's step is
            # set by its launches, and the refresh launch is one of them.
            inner_skin = (INNER_SKIN if num_atoms > _SMALL_SYSTEM and INNER_SKIN < float(skin)
                          else None)
        grid = {}
        if self._npt:
            # wider sub-bins and more halo room: the cell may shrink
            multiplier = float(capacity_multiplier)
            grid = {"grid_margin": NPT_GRID_MARGIN,
                    "halo_margin": multiplier + 4 * NPT_GRID_MARGIN}
        self._neighbors = _MDNeighbors(model.cutoff, skin, capacity_multiplier, self.positions,
                                       cell, types, inner_skin=inner_skin,
                                       reserve=potential_reserve(model), **grid)
        if types is not None:
            # the slot-ordered types the force kernel reads, for both models of a sum
            pot.atom_types = _native.ptr(self._neighbors.buffers["types"])
            if pot.next:
                self._keep[2].atom_types = pot.atom_types
        slots, coef = _native.ctypes.c_int64(), _native.ctypes.c_int64()
        _native.call("md_workspace", _native.dtype_tag(dtype), _native.ctypes.byref(pot),
                     _native.ctypes.c_int64(num_atoms), _native.ctypes.byref(slots),
                     _native.ctypes.byref(coef))
        self.coefficients = torch.zeros(max(coef.value, 256), dtype=torch.uint8, device=device)
        self.partials = torch.zeros(8 * 8, **wide)
        self._system = MDSystem(
            num_atoms=num_atoms, positions=_native.ptr(self.positions),
            velocities=_native.ptr(self.velocities), forces=_native.ptr(self.forces),
            images=_native.ptr(self.images), masses=_native.ptr(self.slot_masses),
            inverse_masses=_native.ptr(self.inverse_masses), cell=_native.ptr(self.cell),
            energy=_native.ptr(self.energy.view(1)), virial=_native.ptr(self.virial),
            thermostat=_native.ptr(self.thermostat), chain_masses=_native.ptr(self.chain_masses),
            kinetic=_native.ptr(self.kinetic), partials=_native.ptr(self.partials),
            partial_slots=8, coefficients=_native.ptr(self.coefficients),
            coefficient_bytes=self.coefficients.numel(), barostat=_native.ptr(self.barostat),
        )
        self._settings = MDSettings(
            timestep=timestep, kT=self._kT, degrees_of_freedom=self._degrees_of_freedom,
            thermostat=1 if self._nvt else 0, virial=1 if self._virial_every_step else 0,
            pressure=self._pressure, box_mass=self._box_mass, alpha=self._alpha,
        )
        self._groups = _GROUPS
        if force_group is not None and force_group not in self._groups:
            raise ValueError(f"force_group must be one of {self._groups}")
        if force_group is not None:
            # a fixed schedule: the force sums' order, hence every bit, independent of timing
            self._neighbors.native.group = force_group
        self._first_build()
        if self._neighbors.inner_skin is not None:
            _native.call("md_refresh", _native.dtype_tag(dtype), _native.ctypes.byref(self._system),
                         _native.ctypes.byref(pot), _native.ctypes.byref(self._neighbors.native),
                         _native.stream(device))
        self.group = self._tune_group() if force_group is None else force_group
        self._evaluate(virial=True)
        self.kinetic.copy_(0.5 * (self.slot_masses.double()[:, None]
                                  * self.velocities.double().square()).sum())
        self._failure: str | None = None
        # without graphs (a simulation stepped in a batch) each step is launched as it comes
        self._graph = self._capture() if graphs else None
        self._chunk_graph = self._capture(_CHUNK_STEPS) if graphs else None

    def _first_build(self) -> None:
        """Sort and count (rows and bricks' halos), size them, then fill. Bricks shrink when the
        largest halo would not fit in LDS."""
        nb = self._neighbors
        shapes = brick_shapes(nb.dims, self._num_atoms, self._dtype, nb.reserve)
        margin = nb.halo_margin
        for k, shape in enumerate(shapes):
            nb.set_bricks(shape)
            self._set_partials(nb.bricks)
            nb.buffers["scan_temp"][656:660].zero_()
            self._rebuild(forced=True)
            nb._check_status()
            needed = nb.halo_needed
            if margin * needed <= halo_limit(self._dtype) - nb.reserve or k == len(shapes) - 1:
                break
        if needed > halo_limit(self._dtype):
            raise ValueError(
                f"a brick's halo holds {needed} atoms, more than the {halo_limit(self._dtype)} "
                "the force kernel's LDS takes (the system is too dense for its sub-bins)")
        nb.size_rows(int(nb.buffers["counts"].max()))
        nb.size_halo(needed)
        self._rebuild(forced=True)
        nb._check_status()
        nb.rebuild_count.fill_(1)

    def _set_partials(self, bricks: int) -> None:
        """One fp64 partial slot (8 sums) per force block: a block per brick, rounded up to a
        multiple of 8."""
        slots = max(8, -(-bricks // 8) * 8)
        self.partials = torch.zeros(slots * 8, dtype=torch.float64, device=self._device)
        self._system.partials = _native.ptr(self.partials)
        self._system.partial_slots = slots

    # -- the step ---------------------------------------------------------------------------

    def _rebuild(self, forced: bool) -> None:
        _native.call("md_rebuild", _native.dtype_tag(self._dtype), _native.ctypes.byref(self._system),
                     _native.ctypes.byref(self._neighbors.state),
                     _native.ctypes.byref(self._neighbors.native), 1 if forced else 0,
                     _native.stream(self._device))

    def _launch_step(self) -> None:
        _native.call(
            "md_step", _native.dtype_tag(self._dtype), _native.ctypes.byref(self._system),
            _native.ctypes.byref(self._settings), _native.ctypes.byref(self._pot),
            _native.ctypes.byref(self._neighbors.state),
            _native.ctypes.byref(self._neighbors.native), _native.stream(self._device),
        )

    def _evaluate(self, virial: bool, kick: bool = False) -> None:
        _native.call(
            "md_evaluate", _native.dtype_tag(self._dtype), _native.ctypes.byref(self._system),
            _native.ctypes.byref(self._pot), _native.ctypes.byref(self._neighbors.native),
            1 if virial else 0, 1 if kick else 0, _native.ctypes.c_double(self._timestep),
            _native.stream(self._device),
        )

    def _tune_group(self, launches: int = 3) -> int:
        """The autotuner of the force schedule: lanes per neighbor row (4, 8 or 16), each timed
        with HIP events on this frame; the fastest wins. Syncs."""
        best, best_time = 8, math.inf
        for group in self._groups:
            self._neighbors.native.group = group
            self._evaluate(virial=False)
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(launches):
                self._evaluate(virial=False)
            end.record()
            torch.cuda.synchronize(self._device)
            elapsed = start.elapsed_time(end)
            if elapsed < best_time:
                best, best_time = group, elapsed
        self._neighbors.native.group = best
        self.tuning = {"group": best, "ms_per_evaluation": best_time / launches}
        return best

    def _settle(self) -> None:
        _native.call("md_settle", _native.dtype_tag(self._dtype),
                     _native.ctypes.byref(self._system), _native.stream(self._device))

    def _capture(self, steps: int = 1) -> torch.cuda.CUDAGraph:
        """A graph of `steps` steps: the first with its prologue, the others opened by the step
        before them (each launch reads the settings when it is recorded)."""
        graph = torch.cuda.CUDAGraph()
        torch.cuda.synchronize(self._device)
        with torch.cuda.graph(graph):
            for k in range(steps):
                self._settings.continues = 1 if k > 0 else 0
                self._launch_step()
        self._settings.continues = 0
        torch.cuda.synchronize(self._device)
        return graph

    def _replay(self, steps: int) -> None:
        """Queues `steps` steps on the current stream: whole chunks through the multi-step
        graph, the rest one step at a time."""
        if self._graph is None:
            for _ in range(steps):
                self._launch_step()
            return
        chunks, rest = divmod(steps, _CHUNK_STEPS)
        for _ in range(chunks):
            self._chunk_graph.replay()
        for _ in range(rest):
            self._graph.replay()

    def _to_atoms(self, tensor: Tensor) -> Tensor:
        """A slot-ordered per-atom tensor in atom order."""
        out = torch.empty_like(tensor)
        width = tensor[0].numel() if tensor.dim() > 1 else 1
        _native.call("md_to_atom_order", tensor.element_size(), _native.ptr(tensor),
                     _native.ptr(self._neighbors.buffers["ids"]),
                     _native.ctypes.c_int64(self._num_atoms), width, _native.ptr(out),
                     _native.stream(self._device))
        return out

    def run(self, steps: Integral, *, sample_every: Integral | None = None, echo=False,
            sample_file=None):
        """Advance ``steps`` steps (``md/_simulation.py:198-272``).

        With ``sample_every``, returns the panel at each step that is a multiple of it, as host
        tensors under ``FIELDS`` (and ``conserved_energy`` for NVT) plus ``step``. ``echo``
        prints each sample. HDF5 sample files are not supported (h5py is not part of this
        build). Reads the neighbor status every 1,000 steps and at the end.
        """
        steps = _count("steps", steps)
        if sample_file is not None:
            raise NotImplementedError("HDF5 sample files are not supported")
        if sample_every is not None and _count("sample_every", sample_every) == 0:
            raise ValueError("sample_every must be positive")
        if echo is not False and sample_every is None:
            raise ValueError("echo needs sample_every")
        self._require()
        start = self._step
        rows, due = [], []
        stream = sys.stdout if echo is True else echo
        done = 0
        while done < steps:
            # replay up to the next event: a sample, a status read, or the end
            segment = min(steps - done, _STATUS_INTERVAL - done % _STATUS_INTERVAL)
            if sample_every is not None:
                segment = min(segment, sample_every - (start + done) % sample_every)
            self._replay(segment)
            done += segment
            step = start + done
            if sample_every is not None and step % sample_every == 0:
                # the row stays on the device; only an echo reads it at once
                rows.append(self._panel_row())
                due.append(step)
                if stream:
                    stream.write(f"{step:>10d} " + " ".join(f"{v:14.6f}" for v in rows[-1].tolist())
                                 + "\n")
            if done % _STATUS_INTERVAL == 0 or done == steps:
                self._check(step)
        self._step += steps
        if sample_every is None:
            return None
        table = (torch.stack(rows).cpu() if rows
                 else torch.zeros((0, len(self.fields)), dtype=torch.float64))
        samples = {"step": torch.tensor(due, dtype=torch.int64)}
        samples.update({name: table[:, k] for k, name in enumerate(self.fields)})
        return samples

    def advance(self, steps: Integral) -> None:
        """Queues ``steps`` graph replays on the current stream and returns at once: no host
        sync and no status read (``check`` reads it). For schedulers that keep several
        simulations in flight on their own streams."""
        steps = _count("steps", steps)
        self._require()
        self._replay(steps)
        self._step += steps

    def check(self) -> None:
        """Reads the neighbor status (syncs) and raises ValueError if a step failed."""
        self._check(self._step)

    @property
    def fields(self) -> tuple[str, ...]:
        if self._npt:
            return FIELDS + ("volume", "conserved_energy")
        return FIELDS + (("conserved_energy",) if self._nvt else ())

    @property
    def model(self) -> Potential:
        return self._model

    @property
    def masses(self) -> Tensor:
        return self._masses.clone()

    def sample(self) -> dict[str, float]:
        """The current frame's panel (potential, kinetic, total energy; T; P; NVT: conserved
        energy; NPT: volume). Syncs with the host."""
        return dict(zip(self.fields, self._panel_row().tolist()))

    def _panel_row(self) -> Tensor:
        """The current frame's panel as a float64 device tensor in ``fields`` order, queued on
        the current stream with no host sync (``md/_simulation.py:366-402`` keeps the panel on
        the device the same way)."""
        self._require()
        self._settle()
        self._evaluate(virial=True)
        masses = self.slot_masses.double()
        kinetic = 0.5 * (masses[:, None] * self.velocities.double().square()).sum()
        energy = self.energy.double()
        volume = self.cell.double().prod()
        values = [
            energy,
            kinetic,
            energy + kinetic,
            2.0 * kinetic / (self._degrees_of_freedom * BOLTZMANN),
            (2.0 * kinetic + self.virial.double().trace()) / (3.0 * volume),
        ]
        if self._nvt:
            q, p = self.thermostat[0:3].double(), self.thermostat[3:6].double()
            bath = ((p.square() / (2.0 * self.chain_masses)).sum()
                    + self._degrees_of_freedom * self._kT * q[0] + self._kT * q[1:].sum())
            conserved = energy + kinetic + bath
        if self._npt:
            # md/_ensembles/_npt.py:205-230: plus the barostat's chain (one degree of freedom),
            # P V of the master cell and the box's kinetic energy
            b = self.barostat.double()
            box_q, box_p = b[_BOX_Q:_BOX_Q + 3], b[_BOX_PC:_BOX_PC + 3]
            box_masses = b[_BOX_MASS:_BOX_MASS + 3]
            box_bath = ((box_p.square() / (2.0 * box_masses)).sum()
                        + self._kT * box_q[0] + self._kT * box_q[1:].sum())
            master_volume = b[_MASTER_CELL:_MASTER_CELL + 3].prod()
            values.append(master_volume)
            conserved = conserved + (box_bath + self._pressure * master_volume
                                     + b[_BOX_P].square() / (2.0 * self._box_mass))
        if self._nvt:
            values.append(conserved)
        return torch.stack(values)

    def state(self) -> State:
        """A copy of the current frame (velocities with the pending chain factor applied)."""
        self._require()
        self._settle()
        chain = box = None
        if self._nvt:
            chain = ChainState(self.thermostat[0:3].clone(), self.thermostat[3:6].clone())
        if self._npt:
            b = self.barostat
            box = BarostatState(b[_BOX_P].clone(), ChainState(b[_BOX_Q:_BOX_Q + 3].clone(),
                                                              b[_BOX_PC:_BOX_PC + 3].clone()))
        return State(self._to_atoms(self.positions), self._to_atoms(self.velocities),
                     self._to_atoms(self.forces), self.cell.clone(), self._to_atoms(self.images),
                     self._step, chain, box)

    def _check(self, step: int) -> None:
        try:
            self._neighbors._check_status()
        except ValueError as error:
            self._failure = f"by step {step}: {error}"
            raise ValueError(f"the simulation failed {self._failure}") from error

    def _require(self) -> None:
        if self._failure is not None:
            raise RuntimeError(
                f"the simulation failed {self._failure}; start a new one from an earlier state()",
            )


def nve(model: Potential, positions: Tensor, velocities: Tensor, masses: Tensor, cell: Tensor, *,
        timestep: Real, skin: Real = 0.1, capacity_multiplier: Real = 1.25,
        images: Tensor | None = None, step: Integral = 0) -> Simulation:
    """Constant energy by velocity Verlet (``md/_ensembles/_nve.py:70-107``)."""
    return Simulation(model, positions, velocities, masses, cell, timestep=timestep, skin=skin,
                      capacity_multiplier=capacity_multiplier, images=images, step=step)


def nvt(model: Potential, positions: Tensor, velocities: Tensor, masses: Tensor, cell: Tensor, *,
        timestep: Real, temperature: Real, thermostat_timescale: Real, skin: Real = 0.1,
        capacity_multiplier: Real = 1.25, images: Tensor | None = None, step: Integral = 0,
        thermostat: ChainState | None = None) -> Simulation:
    """Constant temperature by velocity Verlet and a Nose-Hoover chain of 3
    (``md/_ensembles/_nvt.py:132-179``): chain masses g kT tau^2, kT tau^2, kT tau^2."""
    for name, value in (("temperature", temperature), ("thermostat_timescale", thermostat_timescale)):
        _positive(name, value)
    return Simulation(model, positions, velocities, masses, cell, timestep=timestep, skin=skin,
                      capacity_multiplier=capacity_multiplier, images=images, step=step,
                      temperature=temperature, thermostat_timescale=thermostat_timescale,
                      thermostat=thermostat)


def npt(model: Potential, positions: Tensor, velocities: Tensor, masses: Tensor, cell: Tensor, *,
        timestep: Real, temperature: Real, pressure: Real, thermostat_timescale: Real,
        barostat_timescale: Real, skin: Real = 0.1, capacity_multiplier: Real = 1.25,
        images: Tensor | None = None, step: Integral = 0, thermostat: ChainState | None = None,
        barostat: BarostatState | None = None) -> Simulation:
    """Constant pressure and temperature by isotropic Martyna-Tobias-Klein dynamics
    (``md/_ensembles/_npt.py:270-333``): the arguments of ``nvt`` plus ``pressure`` in
    kJ/mol/nm^3 (``units.BAR`` converts), ``barostat_timescale`` tau_b in ps and, to continue a
    run, the ``barostat`` state. Nose-Hoover chains of 3 act on the particles and on the
    barostat; the box mass is 3 (N + 1) kT tau_b^2; the cell keeps its aspect ratio and scales
    about the origin. The panel adds ``volume`` (nm^3) and ``conserved_energy``. The step's
    sub-bins are chosen ``NPT_GRID_MARGIN`` wider than the list needs, so the cell may shrink by
    that fraction of its first length; further, the run stops with ``ValueError``."""
    return Simulation(model, positions, velocities, masses, cell, timestep=timestep, skin=skin,
                      capacity_multiplier=capacity_multiplier, images=images, step=step,
                      temperature=temperature, thermostat_timescale=thermostat_timescale,
                      thermostat=thermostat, pressure=pressure,
                      barostat_timescale=barostat_timescale, barostat=barostat)


__all__ = ("FIELDS", "Simulation", "npt", "nve", "nvt")
