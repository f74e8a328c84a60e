"""The neighbor list: pairs of atoms closer than cutoff + skin, kept on the GPU.

Host logic of the reference's ``neighbors/_neighbor_list.py`` (validation, grid dims, the
all-pairs choice, capacity sizing, buffer ownership, status decoding), over this library's
gfx950 operators. With ``full=True`` (an extension, used by the MD step and the batched path)
the list also keeps every pair in both orientations, sliced-ELL by rows of 64, so a fused force
kernel sums each atom's forces in registers; the reference's half list ``pairs`` is built in
the same pass and stays bit-exact.
"""

import math
from dataclasses import dataclass, fields
from numbers import Real

import torch
from torch import Tensor

from .. import _native, space
from . import _ops
from ._pairs import NeighborPairs

_INT32_MAX = 2**31 - 1
_INVALID_CELL, _NON_FINITE, _OVERFLOW, _FULL_OVERFLOW = 1, 2, 8, 16
_SLICE = 64


@dataclass
class _Buffers:
    """The device tensors a build writes, in the operators' argument order."""

    dims: Tensor
    widths: Tensor
    atom_bin: Tensor
    bin_counts: Tensor
    bin_offsets: Tensor
    unordered_ids: Tensor
    sorted_ids: Tensor
    sorted_positions: Tensor
    scan_temp: Tensor
    pair_counts: Tensor
    pair_offsets: Tensor
    pair_total: Tensor
    status: Tensor
    pairs: Tensor
    bits: Tensor

    def args(self) -> tuple[Tensor, ...]:
        """The tensors themselves, in argument order."""
        return tuple(getattr(self, field.name) for field in fields(self))

    def named(self) -> dict[str, Tensor]:
        return {field.name: getattr(self, field.name) for field in fields(self)}


@dataclass
class _Reuse:
    """The reuse check's tensors: the frame of the last rebuild, the first grid's bins, the
    threshold and scales, the rebuild flag and count."""

    reference_positions: Tensor
    reference_cell: Tensor
    limits: Tensor
    parameters: Tensor
    rebuild: Tensor
    rebuild_count: Tensor


def _check_settings(cutoff, skin, capacity_multiplier) -> None:
    """``_neighbor_list.py:59-73``."""
    for name, value in (("cutoff", cutoff), ("skin", skin), ("capacity_multiplier", capacity_multiplier)):
        if isinstance(value, bool) or not isinstance(value, Real) or not math.isfinite(value):
            raise ValueError(f"{name} must be a finite number")
    if cutoff <= 0 or skin < 0 or capacity_multiplier < 1:
        raise ValueError(
            "cutoff must be positive, skin nonnegative and capacity_multiplier at least 1",
        )


def _check_inner_cutoffs(inner_cutoffs, cutoff) -> None:
    """``_neighbor_list.py:76-84``."""
    if not isinstance(inner_cutoffs, (tuple, list)):
        raise ValueError("inner_cutoffs must be a tuple of cutoffs in nm")
    for value in inner_cutoffs:
        if isinstance(value, bool) or not isinstance(value, Real) or not 0 < value < cutoff:
            raise ValueError("every inner cutoff must be a number between 0 and the cutoff")
    if len(set(inner_cutoffs)) != len(inner_cutoffs):
        raise ValueError("the inner cutoffs must be distinct")


def _check_frame(positions: Tensor, cell: Tensor) -> None:
    """``_neighbor_list.py:87-100``."""
    for name, value in (("positions", positions), ("cell", cell)):
        if not isinstance(value, Tensor) or value.layout is not torch.strided:
            raise TypeError(f"{name} must be a dense tensor")
    if positions.ndim != 2 or positions.shape[1] != 3 or not 1 <= positions.shape[0] <= _INT32_MAX:
        raise ValueError("positions must have shape [N, 3] with 1 <= N < 2**31")
    if positions.dtype not in (torch.float32, torch.float64) or not positions.is_cuda:
        raise TypeError("positions must be a CUDA float32 or float64 tensor")
    if cell.shape != (3,) or cell.dtype != positions.dtype or cell.device != positions.device:
        raise ValueError("cell must be a [3] tensor with the positions' dtype and device")


def _check_cell_lengths(lengths: list[float], cutoff: float) -> None:
    if not all(math.isfinite(length) and length > 2 * cutoff for length in lengths):
        raise ValueError("every cell length must be finite and above twice the cutoff")


def _check_finite(positions: Tensor) -> None:
    if not bool(torch.isfinite(positions).all()):
        raise ValueError("positions must be finite")


def grid_dims(lengths: list[float], radius: float, dtype: torch.dtype) -> list[int]:
    """Bins per axis, each wider than the list radius plus a 4 eps L margin for the rounding of
    binning (``_neighbor_list.py:115-122``)."""
    eps = torch.finfo(dtype).eps
    return [max(1, math.floor(length / (radius + 4 * eps * length))) for length in lengths]


def _check_bin_count(num_bins: int) -> None:
    if num_bins >= _INT32_MAX:
        raise ValueError("the grid must have fewer than 2**31 - 1 bins")


def _all_pairs_bits(num_atoms: int, dims: list[int], device) -> Tensor:
    """Bit rows of the all-pairs path, chosen for good when the first grid has at most three bins
    per axis (``_neighbor_list.py:131-140``); empty on the grid path."""
    if max(dims) > 3:
        return torch.empty(0, dtype=torch.int32, device=device)
    return torch.zeros(num_atoms * ((num_atoms + 31) // 32), dtype=torch.int32, device=device)


def _check_capacity(capacity: int, count: int) -> None:
    if capacity > _INT32_MAX:
        raise ValueError(
            f"the pair capacity {capacity}, capacity_multiplier times {count} pairs, "
            "exceeds the int32 range",
        )


def _check_allocated_like(positions: Tensor, allocated: Tensor) -> None:
    if (
        positions.shape != allocated.shape
        or positions.dtype != allocated.dtype
        or positions.device != allocated.device
    ):
        raise ValueError("positions must match the allocated shape, dtype and device")


class NeighborList:
    """Pairs of atoms closer than the list radius, cutoff + skin, in nm.

    ``allocate`` sizes the list to ``capacity_multiplier`` times a first frame's pair count and
    builds it; ``update`` reuses the pairs until an atom has moved far enough that a pair within
    the cutoff could be missing, then rebuilds; ``tensors()`` returns the pairs. Cells are
    orthorhombic and periodic; positions may lie outside the cell. ``inner_cutoffs`` keeps inner
    lists at smaller cutoffs, as the reference does.

    Raises ``ValueError`` unless ``cutoff`` is positive, ``skin`` nonnegative and
    ``capacity_multiplier`` at least 1, all finite, and the inner cutoffs distinct numbers
    between 0 and ``cutoff``.
    """

    def __init__(
        self,
        cutoff: Real,
        *,
        skin: Real = 0.0,
        capacity_multiplier: Real = 1.25,
        inner_cutoffs: tuple[Real, ...] = (),
        full: bool = False,
    ):
        _check_settings(cutoff, skin, capacity_multiplier)
        _check_inner_cutoffs(inner_cutoffs, cutoff)
        self._cutoff = float(cutoff)
        self._skin = float(skin)
        self._capacity_multiplier = float(capacity_multiplier)
        self._full_requested = bool(full)
        self._inner = tuple(
            NeighborList(value, skin=skin, capacity_multiplier=capacity_multiplier, full=full)
            for value in sorted(inner_cutoffs)
        )
        self._buffers: _Buffers | None = None
        self._reuse: _Reuse | None = None
        self._full: dict[str, Tensor] | None = None
        self._native = None
        self._failure: str | None = None

    @property
    def cutoff(self) -> float:
        return self._cutoff

    @property
    def radius(self) -> float:
        """The list radius, cutoff + skin, in nm."""
        return self._cutoff + self._skin

    @property
    def all_pairs(self) -> bool:
        """Whether the list tests all pairs (a first grid of at most three bins per axis)."""
        return self._buffers is not None and self._buffers.bits.numel() > 0

    def allocate(self, positions: Tensor, cell: Tensor) -> None:
        """Count the pairs of positions [N, 3] in the cell [3], size the list and build it.

        Syncs with the host. Raises ``TypeError`` or ``ValueError`` for invalid positions or
        cell, and ``ValueError`` when the grid or the pairs do not fit int32.
        """
        _check_frame(positions, cell)
        lengths = [float(length) for length in cell.tolist()]
        _check_cell_lengths(lengths, self._cutoff)
        _check_finite(positions)
        dims = grid_dims(lengths, self.radius, positions.dtype)
        num_atoms, num_bins = positions.shape[0], math.prod(dims)
        _check_bin_count(num_bins)
        device, dtype = positions.device, positions.dtype
        integers = {"dtype": torch.int32, "device": device}
        scan_bytes = max(
            _ops.scan_temp_bytes(positions, num_bins + 1),
            _ops.scan_temp_bytes(positions, num_atoms),
        )
        buffers = _Buffers(
            dims=torch.tensor(dims, **integers),
            widths=torch.tensor(
                [length / count for length, count in zip(lengths, dims)], dtype=dtype, device=device
            ),
            atom_bin=torch.empty(num_atoms, **integers),
            bin_counts=torch.zeros(num_bins + 1, **integers),
            bin_offsets=torch.empty(num_bins + 1, **integers),
            unordered_ids=torch.empty(num_atoms, **integers),
            sorted_ids=torch.empty(num_atoms, **integers),
            sorted_positions=torch.empty((num_atoms, 3), dtype=dtype, device=device),
            scan_temp=torch.zeros(scan_bytes, dtype=torch.uint8, device=device),
            pair_counts=torch.empty(num_atoms, **integers),
            pair_offsets=torch.empty(num_atoms, **integers),
            pair_total=torch.zeros(1, **integers),
            status=torch.zeros(2, **integers),
            pairs=torch.empty((2, 1), **integers),  # sized after the count
            bits=_all_pairs_bits(num_atoms, dims, device),
        )
        full = None
        if self._full_requested:
            slices = (num_atoms + _SLICE - 1) // _SLICE
            full = {
                "counts": torch.zeros(num_atoms, **integers),
                "slices": torch.zeros(slices + 1, **integers),
                "index": torch.zeros(1, **integers),  # sized after the count
            }
        cell = cell.detach().contiguous()
        wrapped, _ = space.wrap(positions.detach(), cell)
        _ops.count_first_frame(wrapped, cell, self.radius, buffers.named(), full)
        if full is None:
            count = int(buffers.pair_counts.sum(dtype=torch.int64))
        else:
            count, full_count = (
                int(v)
                for v in torch.stack(
                    [buffers.pair_counts.sum(dtype=torch.int64), full["slices"][-1].to(torch.int64)]
                ).tolist()
            )
        capacity = max(8, math.ceil(self._capacity_multiplier * count))
        _check_capacity(capacity, count)
        buffers.pairs = torch.empty((2, capacity), **integers)
        if full is not None:
            full_capacity = max(_SLICE, math.ceil(self._capacity_multiplier * full_count))
            _check_capacity(full_capacity, full_count)
            full["index"] = torch.zeros(full_capacity, **integers)
        _ops.fill_first_frame(wrapped, cell, self.radius, buffers.named(), full)
        self._reuse = _Reuse(
            reference_positions=wrapped.clone(),
            reference_cell=cell.clone(),
            limits=torch.tensor(dims, **integers),
            parameters=torch.zeros(4, dtype=dtype, device=device),
            rebuild=torch.zeros(1, **integers),
            rebuild_count=torch.ones(1, **integers),
        )
        self._buffers = buffers
        self._full = full
        self._native = _ops.native_buffers(wrapped, buffers.named(), full)
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
        buffers = self._live_buffers()
        _check_frame(positions, cell)
        _check_allocated_like(positions, buffers.sorted_positions)
        wrapped, _ = space.wrap(positions.detach(), cell.detach())
        self._update_wrapped(wrapped, cell.detach().contiguous())
        self._check_status()

    def _update_wrapped(self, positions: Tensor, cell: Tensor) -> None:
        """Reuse or rebuild the pairs, and the inner lists', for contiguous positions in [0, L).

        No checks, no allocation and no host sync, so a HIP graph can capture it.
        """
        reuse = self._reuse
        native = self._native
        _native.call(
            "update",
            _native.dtype_tag(positions.dtype),
            _native.ptr(positions),
            _native.ptr(cell),
            _native.ctypes.c_double(self._cutoff),
            _native.ctypes.c_double(self.radius),
            _native.ptr(reuse.reference_positions),
            _native.ptr(reuse.reference_cell),
            _native.ptr(reuse.limits),
            _native.ptr(reuse.parameters),
            _native.ptr(reuse.rebuild),
            _native.ptr(reuse.rebuild_count),
            _native.ctypes.byref(native),
            _native.stream(positions.device),
        )
        for inner in self._inner:
            inner._update_wrapped(positions, cell)

    def _rebuild_wrapped(self, positions: Tensor, cell: Tensor) -> None:
        """Rebuild for contiguous positions in [0, L) whatever the reuse rule: an infinite
        reference cell forces it (``_neighbor_list.py:326-335``). No host sync."""
        for listed in (self, *self._inner):
            listed._reuse.reference_cell.fill_(math.inf)
        self._update_wrapped(positions, cell)

    def _check_status(self) -> None:
        """Raise ``ValueError`` and fail the list if an update of it or an inner list failed.

        Syncs with the host (``_neighbor_list.py:337-361``).
        """
        flags, needed = self._live_buffers().status.tolist()
        if flags:
            causes = []
            if flags & _INVALID_CELL:
                causes.append("every cell length must be finite and above twice the cutoff")
            if flags & _NON_FINITE:
                causes.append("positions must be finite")
            if flags & _OVERFLOW:
                causes.append(
                    f"the pair capacity {self.pair_capacity} is too small: the frame has "
                    f"{needed} pairs within {self.radius} nm; raise capacity_multiplier",
                )
            if flags & _FULL_OVERFLOW:
                causes.append(
                    f"the full list's capacity {self.full_capacity} is too small for this frame; "
                    "raise capacity_multiplier",
                )
            self._failure = "; ".join(causes)
            raise ValueError(self._failure)
        try:
            for inner in self._inner:
                inner._check_status()
        except ValueError as error:
            self._failure = str(error)
            raise

    def tensors(self) -> NeighborPairs:
        """The pairs of the last build, as the ``NeighborPairs`` that models take."""
        buffers = self._live_buffers()
        inner = tuple(listed.tensors() for listed in self._inner)
        return NeighborPairs(buffers.pairs, buffers.sorted_ids.numel(), self._cutoff, inner)

    @property
    def pair_capacity(self) -> int:
        """C, the columns of the list's own pair index; 0 before ``allocate``."""
        return self._buffers.pairs.shape[1] if self._buffers is not None else 0

    @property
    def full_capacity(self) -> int:
        """Entries of the full list; 0 without one."""
        return self._full["index"].numel() if self._full is not None else 0

    @property
    def num_pairs(self) -> int:
        """The number of the list's own pairs at the last build. Syncs with the host."""
        return int(self._live_buffers().pair_total.item())

    @property
    def _rebuild_count(self) -> int:
        """The builds since allocate, counting allocate's own. Syncs with the host."""
        return int(self._reuse.rebuild_count.item())

    def _max_moved(self) -> float:
        """The largest squared displacement the last reuse check measured, in nm^2: the maximum
        of the 64 per-block slots after the first 64 bytes of scan_temp. Syncs with the host."""
        import struct

        raw = int(self._buffers.scan_temp[64:64 + 8 * 64].view(torch.int64).max().item())
        if self._buffers.sorted_positions.dtype == torch.float32:
            return struct.unpack("<f", struct.pack("<I", raw & 0xFFFFFFFF))[0]
        return struct.unpack("<d", struct.pack("<q", raw))[0]

    def _live_buffers(self) -> _Buffers:
        if self._buffers is None:
            raise RuntimeError("call allocate() first")
        if self._failure is not None:
            raise RuntimeError(
                f"the neighbor list failed ({self._failure}); call allocate() again",
            )
        return self._buffers


__all__ = ("NeighborList",)
