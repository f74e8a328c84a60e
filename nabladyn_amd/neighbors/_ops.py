"""The ``nabladyn_neighbors`` operators, registered with ``torch.library`` under the reference's
exact schemas (``neighbors/csrc/ops.cpp:11-41``), each implemented by one C-ABI call.

``scan_temp_bytes`` is registered for ``CompositeExplicitAutograd`` and the three kernel ops
for ``CUDA``, which on PyTorch-ROCm is the dispatch key of HIP devices. Argument errors raise
``RuntimeError`` as the reference's ``TORCH_CHECK`` does (``runtime/entry_points.cu:24-103``);
errors in the data set the list's status bits and never raise here.
"""

import torch

from .. import _native

_BUFFERS = (
    "Tensor(a!) dims, Tensor(b!) widths, Tensor(c!) atom_bin, Tensor(d!) bin_counts, "
    "Tensor(e!) bin_offsets, Tensor(f!) unordered_ids, Tensor(g!) sorted_ids, "
    "Tensor(h!) sorted_positions, Tensor(i!) scan_temp, Tensor(j!) pair_counts, "
    "Tensor(k!) pair_offsets, Tensor(l!) pair_total, Tensor(m!) status, Tensor(n!) pairs, "
    "Tensor(t!) bits"
)
SCHEMAS = {
    "scan_temp_bytes": "scan_temp_bytes(Tensor reference, int items) -> int",
    "count_first_frame": f"count_first_frame(Tensor positions, Tensor cell, float radius, {_BUFFERS}) -> ()",
    "fill_first_frame": f"fill_first_frame(Tensor positions, Tensor cell, float radius, {_BUFFERS}) -> ()",
    "update": (
        "update(Tensor positions, Tensor cell, float cutoff, float radius, "
        "Tensor(o!) reference_positions, Tensor(p!) reference_cell, Tensor limits, "
        f"Tensor(q!) reuse, Tensor(r!) rebuild, Tensor(s!) rebuild_count, {_BUFFERS}) -> ()"
    ),
}
BUFFER_NAMES = (
    "dims", "widths", "atom_bin", "bin_counts", "bin_offsets", "unordered_ids", "sorted_ids",
    "sorted_positions", "scan_temp", "pair_counts", "pair_offsets", "pair_total", "status",
    "pairs", "bits",
)


def _check(tensor: torch.Tensor, reference: torch.Tensor, dtype: torch.dtype, name: str) -> None:
    """``entry_points.cu:24-38``."""
    if not (tensor.is_cuda and tensor.device == reference.device):
        raise RuntimeError(f"{name} must be on the positions' CUDA device")
    if not tensor.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")
    if tensor.dtype != dtype:
        raise RuntimeError(f"{name} has the wrong dtype")


def check_buffers(positions: torch.Tensor, buffers: dict[str, torch.Tensor]) -> None:
    """``Buffers::check_against`` (``entry_points.cu:54-103``)."""
    dtype = positions.dtype
    if dtype not in (torch.float32, torch.float64):
        raise RuntimeError("positions must be float32 or float64")
    if positions.dim() != 2 or positions.size(1) != 3:
        raise RuntimeError("positions must be [N, 3]")
    n = positions.size(0)
    _check(positions, positions, dtype, "positions")
    _check(buffers["widths"], positions, dtype, "widths")
    _check(buffers["sorted_positions"], positions, dtype, "sorted_positions")
    for name in BUFFER_NAMES:
        if name in ("widths", "sorted_positions", "scan_temp"):
            continue
        _check(buffers[name], positions, torch.int32, name)
    _check(buffers["scan_temp"], positions, torch.uint8, "scan_temp")
    b = buffers
    if not (
        b["dims"].numel() == 3 and b["widths"].numel() == 3 and b["atom_bin"].numel() == n
        and b["unordered_ids"].numel() == n and b["sorted_ids"].numel() == n
        and b["sorted_positions"].shape == positions.shape and b["bin_counts"].numel() >= 2
        and b["bin_offsets"].numel() == b["bin_counts"].numel() and b["pair_counts"].numel() == n
        and b["pair_offsets"].numel() == n and b["pair_total"].numel() == 1
        and b["status"].numel() == 2
    ):
        raise RuntimeError("the list's buffers do not match the positions")
    pairs = b["pairs"]
    if not (pairs.dim() == 2 and pairs.size(0) == 2 and 1 <= pairs.size(1) <= 2**31 - 1):
        raise RuntimeError("pairs must be [2, C] with int32 C")
    words = n * ((n + 31) // 32)
    if not (b["bits"].numel() == 0 or b["bits"].numel() == words):
        raise RuntimeError("bits must be empty or hold N rows of ceil(N / 32) words")


def native_buffers(
    positions: torch.Tensor,
    buffers: dict[str, torch.Tensor],
    full: dict[str, torch.Tensor] | None = None,
) -> _native.NeighborBuffers:
    """The C-ABI struct of a list's buffers (and its full list, when given)."""
    b = buffers
    s = _native.NeighborBuffers(
        **{name: _native.ptr(b[name]) for name in BUFFER_NAMES},
        num_atoms=positions.size(0),
        bin_slots=b["bin_counts"].numel(),
        scan_temp_bytes=b["scan_temp"].numel(),
        capacity=b["pairs"].size(1),
        bits_words=b["bits"].numel(),
    )
    if full is not None:
        s.full_counts = _native.ptr(full["counts"])
        s.full_slices = _native.ptr(full["slices"])
        s.full_index = _native.ptr(full["index"])
        s.full_capacity = full["index"].numel()
    return s


def scan_temp_bytes(reference: torch.Tensor, items: int) -> int:
    out = _native.ctypes.c_int64()
    _native.call("scan_temp_bytes", _native.ctypes.c_int64(int(items)), _native.ctypes.byref(out))
    return int(out.value)


def _first_frame(name, positions, cell, radius, buffers, full=None):
    check_buffers(positions, buffers)
    _check(cell, positions, positions.dtype, "cell")
    native = native_buffers(positions, buffers, full)
    with torch.cuda.device(positions.device):
        _native.call(
            name,
            _native.dtype_tag(positions.dtype),
            _native.ptr(positions),
            _native.ptr(cell),
            _native.ctypes.c_double(float(radius)),
            _native.ctypes.byref(native),
            _native.stream(positions.device),
        )


def count_first_frame(positions, cell, radius, buffers, full=None) -> None:
    _first_frame("count_first_frame", positions, cell, radius, buffers, full)


def fill_first_frame(positions, cell, radius, buffers, full=None) -> None:
    _first_frame("fill_first_frame", positions, cell, radius, buffers, full)


def update(positions, cell, cutoff, radius, reference_positions, reference_cell, limits, reuse,
           rebuild, rebuild_count, buffers, full=None, checked: bool = True) -> None:
    """``entry_points.cu:401-508``: checks (unless ``checked`` is False), then one C-ABI call."""
    dtype = positions.dtype
    if checked:
        check_buffers(positions, buffers)
        _check(cell, positions, dtype, "cell")
        _check(reference_positions, positions, dtype, "reference_positions")
        _check(reference_cell, positions, dtype, "reference_cell")
        _check(limits, positions, torch.int32, "limits")
        _check(reuse, positions, dtype, "reuse")
        _check(rebuild, positions, torch.int32, "rebuild")
        _check(rebuild_count, positions, torch.int32, "rebuild_count")
        if not (
            cell.numel() == 3 and reference_cell.numel() == 3 and limits.numel() == 3
            and reference_positions.shape == positions.shape and reuse.numel() == 4
            and rebuild.numel() == 1 and rebuild_count.numel() == 1
        ):
            raise RuntimeError("the list's update buffers do not match the positions")
    native = native_buffers(positions, buffers, full)
    _native.call(
        "update",
        _native.dtype_tag(dtype),
        _native.ptr(positions),
        _native.ptr(cell),
        _native.ctypes.c_double(float(cutoff)),
        _native.ctypes.c_double(float(radius)),
        _native.ptr(reference_positions),
        _native.ptr(reference_cell),
        _native.ptr(limits),
        _native.ptr(reuse),
        _native.ptr(rebuild),
        _native.ptr(rebuild_count),
        _native.ctypes.byref(native),
        _native.stream(positions.device),
    )


def _register() -> None:
    library = torch.library.Library("nabladyn_neighbors", "DEF")
    for schema in SCHEMAS.values():
        library.define(schema)

    def _scan_temp_bytes(reference, items):
        return scan_temp_bytes(reference, items)

    def _count(positions, cell, radius, *bufs):
        count_first_frame(positions, cell, radius, dict(zip(BUFFER_NAMES, bufs)))

    def _fill(positions, cell, radius, *bufs):
        fill_first_frame(positions, cell, radius, dict(zip(BUFFER_NAMES, bufs)))

    def _update(positions, cell, cutoff, radius, reference_positions, reference_cell, limits,
                reuse, rebuild, rebuild_count, *bufs):
        with torch.cuda.device(positions.device):
            update(positions, cell, cutoff, radius, reference_positions, reference_cell, limits,
                   reuse, rebuild, rebuild_count, dict(zip(BUFFER_NAMES, bufs)))

    library.impl("scan_temp_bytes", _scan_temp_bytes, "CompositeExplicitAutograd")
    library.impl("count_first_frame", _count, "CUDA")
    library.impl("fill_first_frame", _fill, "CUDA")
    library.impl("update", _update, "CUDA")
    return library


_LIBRARY = None


def ops():
    """Return ``torch.ops.nabladyn_neighbors``, registering the operators once per process."""
    global _LIBRARY
    if _LIBRARY is None:
        with _native._LOCK:
            if _LIBRARY is None:
                _LIBRARY = _register()
    return torch.ops.nabladyn_neighbors


__all__ = ("ops", "SCHEMAS")
