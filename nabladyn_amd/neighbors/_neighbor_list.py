"""The neighbor list (``neighbors/_neighbor_list.py``): pairs of atoms closer than cutoff + skin,
kept on the GPU and reused until the reference's reuse rule asks for a rebuild.

A list here is three things:

- ``_Settings``: cutoff, skin, capacity multiplier and whether a full list is kept, checked
  once when the list is made;
- ``_Plan``: what a first frame fixes on the host from its cell alone — the grid, the path
  (all pairs or grid), the scan scratch;
- ``_Device``: every device tensor of the list in one table keyed by the operators' argument
  names (so the C-ABI struct and the ``nabladyn_neighbors`` ops read the same table), the reuse
  rule's tensors, and the optional full list.

Builds and updates are C-ABI calls into the gfx950 library; errors in the data come back as
latched status bits, decoded here into ``ValueError`` through one table of messages. With
``full=True`` (an extension used by the fused force kernels) the build also keeps every pair in
both orientations, sliced by row groups of 64; the reference's half list is built in the same
pass and stays bit-exact.
"""

import math
import struct
from dataclasses import dataclass
from numbers import Real

import torch
from torch import Tensor

from .. import _native, space
from . import _ops
from ._pairs import FullRows, NeighborPairs

_INT32_MAX = 2**31 - 1
_ROWS_PER_SLICE = 64


def _is_number(value) -> bool:
    return isinstance(value, Real) and not isinstance(value, bool) and math.isfinite(value)


@dataclass(frozen=True)
class _Settings:
    cutoff: float
    skin: float
    multiplier: float
    inner: tuple[float, ...]
    full: bool

    @staticmethod
    def checked(cutoff, skin, capacity_multiplier, inner_cutoffs, full) -> "_Settings":
        """The settings, or ``ValueError`` with the reference's messages
        (``_neighbor_list.py:59-84``)."""
        for name, value in (("cutoff", cutoff), ("skin", skin),
                            ("capacity_multiplier", capacity_multiplier)):
            if not _is_number(value):
                raise ValueError(f"{name} must be a finite number")
        if not (cutoff > 0 and skin >= 0 and capacity_multiplier >= 1):
            raise ValueError(
                "cutoff must be positive, skin nonnegative and capacity_multiplier at least 1",
            )
        if not isinstance(inner_cutoffs, (tuple, list)):
            raise ValueError("inner_cutoffs must be a tuple of cutoffs in nm")
        if not all(_is_number(v) and 0 < v < cutoff for v in inner_cutoffs):
            raise ValueError("every inner cutoff must be a number between 0 and the cutoff")
        if len(set(inner_cutoffs)) < len(inner_cutoffs):
            raise ValueError("the inner cutoffs must be distinct")
        return _Settings(float(cutoff), float(skin), float(capacity_multiplier),
                         tuple(sorted(float(v) for v in inner_cutoffs)), bool(full))

    @property
    def radius(self) -> float:
        return self.cutoff + self.skin


def grid_dims(lengths: list[float], radius: float, dtype: torch.dtype) -> list[int]:
    """Bins per axis: the largest count whose bins are wider than the list radius plus a margin
    of 4 eps L for the rounding of binning, at least one (``_neighbor_list.py:115-122``)."""
    eps = torch.finfo(dtype).eps
    return [max(1, math.floor(length / (radius + 4 * eps * length))) for length in lengths]


def pair_capacity_for(count: int, multiplier: float, minimum: int = 8) -> int:
    """The columns for `count` pairs: ``max(8, ceil(multiplier * count))``
    (``_neighbor_list.py:259``); ``ValueError`` past int32."""
    capacity = max(minimum, math.ceil(multiplier * count))
    if capacity > _INT32_MAX:
        raise ValueError(
            f"the pair capacity {capacity}, capacity_multiplier times {count} pairs, "
            "exceeds the int32 range",
        )
    return capacity


def _check_frame(positions, cell) -> None:
    """The frame's types, shapes, dtypes and device (``_neighbor_list.py:87-100``)."""
    if not all(isinstance(v, Tensor) and v.layout is torch.strided for v in (positions, cell)):
        bad = "positions" if not (isinstance(positions, Tensor)
                                  and positions.layout is torch.strided) else "cell"
        raise TypeError(f"{bad} must be a dense tensor")
    n = positions.shape[0] if positions.ndim == 2 else 0
    if positions.ndim != 2 or positions.shape[1] != 3 or not 1 <= n <= _INT32_MAX:
        raise ValueError("positions must have shape [N, 3] with 1 <= N < 2**31")
    if positions.dtype not in (torch.float32, torch.float64) or not positions.is_cuda:
        raise TypeError("positions must be a CUDA float32 or float64 tensor")
    if (cell.shape, cell.dtype, cell.device) != ((3,), positions.dtype, positions.device):
        raise ValueError("cell must be a [3] tensor with the positions' dtype and device")


@dataclass(frozen=True)
class _Plan:
    """What a first frame fixes: the grid, the path and the scratch sizes."""

    lengths: tuple[float, ...]
    dims: tuple[int, ...]
    num_atoms: int
    scan_bytes: int

    @property
    def num_bins(self) -> int:
        return math.prod(self.dims)

    @property
    def all_pairs(self) -> bool:
        # the reference's choice, fixed for the list's life (_neighbor_list.py:131-140)
        return max(self.dims) <= 3

    @staticmethod
    def of(positions: Tensor, cell: Tensor, settings: _Settings) -> "_Plan":
        """Syncs with the host (the cell, and whether every coordinate is finite)."""
        lengths = tuple(float(v) for v in cell.tolist())
        if not all(math.isfinite(v) and v > 2 * settings.cutoff for v in lengths):
            raise ValueError("every cell length must be finite and above twice the cutoff")
        if not bool(torch.isfinite(positions).all()):
            raise ValueError("positions must be finite")
        dims = tuple(grid_dims(list(lengths), settings.radius, positions.dtype))
        n = positions.shape[0]
        if math.prod(dims) >= _INT32_MAX:
            raise ValueError("the grid must have fewer than 2**31 - 1 bins")
        scan = max(_ops.scan_temp_bytes(positions, math.prod(dims) + 1),
                   _ops.scan_temp_bytes(positions, n))
        return _Plan(lengths, dims, n, scan)


class _Device:
    """The list's device tensors: the operators' buffers (one table, in argument order), the
    reuse rule's tensors, and the full list when one is kept."""

    def __init__(self, plan: _Plan, positions: Tensor, settings: _Settings):
        device, dtype, n = positions.device, positions.dtype, plan.num_atoms
        ints = {"dtype": torch.int32, "device": device}
        per_atom = lambda: torch.empty(n, **ints)  # noqa: E731
        bins = plan.num_bins + 1
        words = n * ((n + 31) // 32) if plan.all_pairs else 0
        table = {
            "dims": torch.tensor(plan.dims, **ints),
            "widths": torch.tensor([length / count for length, count in zip(plan.lengths, plan.dims)],
                                   dtype=dtype, device=device),
            "atom_bin": per_atom(),
            "bin_counts": torch.zeros(bins, **ints),
            "bin_offsets": torch.empty(bins, **ints),
            "unordered_ids": per_atom(),
            "sorted_ids": per_atom(),
            "sorted_positions": torch.empty((n, 3), dtype=dtype, device=device),
            "scan_temp": torch.zeros(plan.scan_bytes, dtype=torch.uint8, device=device),
            "pair_counts": per_atom(),
            "pair_offsets": per_atom(),
            "pair_total": torch.zeros(1, **ints),
            "status": torch.zeros(2, **ints),
            "pairs": torch.empty((2, 1), **ints),  # sized once the pairs are counted
            "bits": torch.zeros(words, **ints),
        }
        assert tuple(table) == _ops.BUFFER_NAMES
        self.buffers = table
        self.full = None
        if settings.full:
            self.full = {
                "counts": torch.zeros(n, **ints),
                "slices": torch.zeros((n + _ROWS_PER_SLICE - 1) // _ROWS_PER_SLICE + 1, **ints),
                "index": torch.zeros(1, **ints),  # sized once the pairs are counted
            }
        self.reuse = None
        self.native = None

    def args(self) -> tuple[Tensor, ...]:
        """The buffers in the operators' argument order."""
        return tuple(self.buffers.values())

    def bind(self, positions: Tensor) -> None:
        self.native = _ops.native_buffers(positions, self.buffers, self.full)

    def start_reuse(self, wrapped: Tensor, cell: Tensor, dims) -> None:
        """The reuse rule's state after a first build: its frame, its grid as the limits of
        later grids, the rule's scratch, the flag and the build count (1)."""
        ints = {"dtype": torch.int32, "device": wrapped.device}
        self.reuse = {
            "reference_positions": wrapped.clone(),
            "reference_cell": cell.clone(),
            "limits": torch.tensor(dims, **ints),
            "parameters": torch.zeros(4, dtype=wrapped.dtype, device=wrapped.device),
            "rebuild": torch.zeros(1, **ints),
            "rebuild_count": torch.ones(1, **ints),
        }


def _message_invalid_cell(nl, needed):
    return "every cell length must be finite and above twice the cutoff"


def _message_non_finite(nl, needed):
    return "positions must be finite"


def _message_overflow(nl, needed):
    return (f"the pair capacity {nl.pair_capacity} is too small: the frame has {needed} pairs "
            f"within {nl.radius} nm; raise capacity_multiplier")


def _message_full_overflow(nl, needed):
    return (f"the full list's capacity {nl.full_capacity} is too small for this frame; "
            "raise capacity_multiplier")


# status bit -> the cause it reports (``_neighbor_list.py:337-361``; bit 16 is this library's)
_CAUSES = (
    (_native.STATUS_INVALID_CELL, _message_invalid_cell),
    (_native.STATUS_NON_FINITE, _message_non_finite),
    (_native.STATUS_OVERFLOW, _message_overflow),
    (_native.STATUS_FULL_OVERFLOW, _message_full_overflow),
)


class NeighborList:
    """Pairs of atoms closer than the list radius, cutoff + skin, in nm.

    ``allocate`` sizes the list to ``capacity_multiplier`` times a first frame's pair count and
    builds it; ``update`` reuses the pairs until an atom has moved far enough that a pair within
    the cutoff could be missing, then rebuilds; ``tensors()`` returns the pairs. Cells are
    orthorhombic and periodic; positions may lie outside the cell. ``inner_cutoffs`` keeps inner
    lists at smaller cutoffs, as the reference does. ``full`` (an extension, on by default) also
    keeps every pair in both orientations, which ``quantity`` evaluates with no atomics; the
    half list ``tensors().index`` is the reference's either way. ``full=False`` saves that
    memory (about as much again as the half list).

    Raises ``ValueError`` unless ``cutoff`` is positive, ``skin`` nonnegative and
    ``capacity_multiplier`` at least 1, all finite, and the inner cutoffs distinct numbers
    between 0 and ``cutoff``.
    """

    def __init__(self, cutoff: Real, *, skin: Real = 0.0, capacity_multiplier: Real = 1.25,
                 inner_cutoffs: tuple[Real, ...] = (), full: bool = True):
        self._settings = _Settings.checked(cutoff, skin, capacity_multiplier, inner_cutoffs, full)
        s = self._settings
        self._inner = tuple(NeighborList(value, skin=s.skin, capacity_multiplier=s.multiplier,
                                         full=s.full) for value in s.inner)
        self._device: _Device | None = None
        self._failure: str | None = None

    # ---- properties ---------------------------------------------------------------------------

    @property
    def cutoff(self) -> float:
        return self._settings.cutoff

    @property
    def radius(self) -> float:
        """The list radius, cutoff + skin, in nm."""
        return self._settings.radius

    @property
    def all_pairs(self) -> bool:
        """Whether the list tests all pairs (a first grid of at most three bins per axis)."""
        return self._device is not None and self._device.buffers["bits"].numel() > 0

    @property
    def pair_capacity(self) -> int:
        """C, the columns of the list's own pair index; 0 before ``allocate``."""
        return self._device.buffers["pairs"].shape[1] if self._device is not None else 0

    @property
    def full_capacity(self) -> int:
        """Entries of the full list; 0 without one."""
        full = self._device.full if self._device is not None else None
        return full["index"].numel() if full is not None else 0

    @property
    def num_pairs(self) -> int:
        """The number of the list's own pairs at the last build. Syncs with the host."""
        return int(self._live().buffers["pair_total"].item())

    @property
    def _rebuild_count(self) -> int:
        """The builds since allocate, counting allocate's own. Syncs with the host."""
        return int(self._device.reuse["rebuild_count"].item())

    # ---- building -----------------------------------------------------------------------------

    def allocate(self, positions: Tensor, cell: Tensor) -> None:
        """Count the pairs of positions [N, 3] in the cell [3], size the list and build it.

        Syncs with the host. Raises ``TypeError`` or ``ValueError`` for invalid positions or
        cell, and ``ValueError`` when the grid or the pairs do not fit int32.
        """
        _check_frame(positions, cell)
        settings = self._settings
        plan = _Plan.of(positions, cell, settings)
        dev = _Device(plan, positions, settings)
        cell = cell.detach().contiguous()
        wrapped, _ = space.wrap(positions.detach(), cell)
        # count, size what the count fixes, fill
        _ops.count_first_frame(wrapped, cell, self.radius, dev.buffers, dev.full)
        totals = [dev.buffers["pair_counts"].sum(dtype=torch.int64)]
        if dev.full is not None:
            totals.append(dev.full["slices"][-1].to(torch.int64))
        counts = [int(v) for v in torch.stack(totals).tolist()]
        ints = {"dtype": torch.int32, "device": positions.device}
        dev.buffers["pairs"] = torch.empty(
            (2, pair_capacity_for(counts[0], settings.multiplier)), **ints)
        if dev.full is not None:
            dev.full["index"] = torch.zeros(
                pair_capacity_for(counts[1], settings.multiplier, _ROWS_PER_SLICE), **ints)
        _ops.fill_first_frame(wrapped, cell, self.radius, dev.buffers, dev.full)
        dev.start_reuse(wrapped, cell, plan.dims)
        dev.bind(wrapped)
        self._device = dev
        # the list fails as a whole until every inner list is built
        self._failure = "an inner list was not allocated"
        for inner in self._inner:
            inner.allocate(positions, cell)
        self._failure = None

    def update(self, positions: Tensor, cell: Tensor) -> None:
        """Reuse the pairs for positions [N, 3] in the cell [3], or rebuild them.

        The pairs are reused while no atom has moved more than (R s_min - r_c) / 2 from its
        position at the last rebuild, scaled with the cell (``update/update.cuh:11-19``). Syncs
        with the host. Raises as the reference's ``update`` does (``_neighbor_list.py:281-301``).
        """
        sorted_positions = self._live().buffers["sorted_positions"]
        _check_frame(positions, cell)
        if (positions.shape, positions.dtype, positions.device) != (
                sorted_positions.shape, sorted_positions.dtype, sorted_positions.device):
            raise ValueError("positions must match the allocated shape, dtype and device")
        wrapped, _ = space.wrap(positions.detach(), cell.detach())
        self._update_wrapped(wrapped, cell.detach().contiguous())
        self._check_status()

    def _update_wrapped(self, positions: Tensor, cell: Tensor) -> None:
        """Reuse or rebuild the pairs, and the inner lists', for contiguous positions in [0, L).

        No checks, no allocation and no host sync, so a HIP graph can capture it.
        """
        dev = self._device
        r = dev.reuse
        _native.call(
            "update", _native.dtype_tag(positions.dtype), _native.ptr(positions),
            _native.ptr(cell), _native.ctypes.c_double(self.cutoff),
            _native.ctypes.c_double(self.radius), _native.ptr(r["reference_positions"]),
            _native.ptr(r["reference_cell"]), _native.ptr(r["limits"]),
            _native.ptr(r["parameters"]), _native.ptr(r["rebuild"]),
            _native.ptr(r["rebuild_count"]), _native.ctypes.byref(dev.native),
            _native.stream(positions.device),
        )
        for inner in self._inner:
            inner._update_wrapped(positions, cell)

    def _rebuild_wrapped(self, positions: Tensor, cell: Tensor) -> None:
        """Rebuild for contiguous positions in [0, L) whatever the reuse rule: an infinite
        reference cell forces it (``_neighbor_list.py:326-335``). No host sync."""
        for listed in (self, *self._inner):
            listed._device.reuse["reference_cell"].fill_(math.inf)
        self._update_wrapped(positions, cell)

    # ---- status -------------------------------------------------------------------------------

    def _check_status(self) -> None:
        """Raise ``ValueError`` and fail the list if an update of it or an inner list failed.

        Syncs with the host (``_neighbor_list.py:337-361``).
        """
        flags, needed = self._live().buffers["status"].tolist()
        causes = [message(self, needed) for bit, message in _CAUSES if flags & bit]
        if causes:
            self._failure = "; ".join(causes)
            raise ValueError(self._failure)
        for inner in self._inner:
            try:
                inner._check_status()
            except ValueError as error:
                self._failure = str(error)
                raise

    def _live(self) -> _Device:
        if self._device is None:
            raise RuntimeError("call allocate() first")
        if self._failure is not None:
            raise RuntimeError(f"the neighbor list failed ({self._failure}); call allocate() again")
        return self._device

    # ---- results ------------------------------------------------------------------------------

    def tensors(self) -> NeighborPairs:
        """The pairs of the last build, as the ``NeighborPairs`` that models take."""
        dev = self._live()
        buffers = dev.buffers
        full = None
        if dev.full is not None:
            rows = None if self.all_pairs else buffers["sorted_ids"]
            full = FullRows(rows, dev.full["counts"], dev.full["slices"], dev.full["index"])
        return NeighborPairs(buffers["pairs"], buffers["sorted_ids"].numel(), self.cutoff,
                             tuple(listed.tensors() for listed in self._inner), full)

    def _max_moved(self) -> float:
        """The largest squared displacement the last reuse check measured, in nm^2: the maximum
        of the 64 per-block slots after the first 64 bytes of scan_temp. Syncs with the host."""
        buffers = self._device.buffers
        raw = int(buffers["scan_temp"][64:64 + 8 * 64].view(torch.int64).max().item())
        if buffers["sorted_positions"].dtype == torch.float32:
            return struct.unpack("<f", struct.pack("<I", raw & 0xFFFFFFFF))[0]
        return struct.unpack("<d", struct.pack("<q", raw))[0]


__all__ = ("NeighborList",)
