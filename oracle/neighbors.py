"""ORACLE — test infrastructure, not product: the neighbor list restated on the CPU.

Two independent restatements:

- ``build``/``update_decision``: the C restatement (``neighbors.c``) of the reference's cell
  list, in the list's dtype with no contraction, which reproduces the reference's column order
  (centers in (bin, atom id) order, neighbor bins z, y, x from -1 to +1, candidates ascending;
  all-pairs (i, j) ascending). Reference: ``neighbors/csrc/*`` and ``_neighbor_list.py``.
- ``brute_pairs``: a numpy O(N^2) search, written independently from the formula
  r^2 = (dx dx + dy dy) + dz dz < fl(R) fl(R) with the minimum-image difference of
  ``traversal.cuh:31-38``, evaluated in the dtype (numpy never fuses). It pins the C
  restatement's pair set.

Parity status: pinned. The C build is checked against ``brute_pairs`` on random and lattice
frames in both dtypes, both against the documented counts of ``argon_512.npz``
(``README.md:166-172``), and the reuse rule against its closed forms
(tests/test_oracle_neighbors.py).
"""

import ctypes
import math
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = _HERE / "_build" / "liboracle.so"
_lib = None

INVALID_CELL, NON_FINITE, OVERFLOW = 1, 2, 8


def build_library() -> Path:
    """Compile ``neighbors.c`` with the oracle Makefile; return the library path."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB


def lib():
    """Load the compiled oracle, building it first if it is missing."""
    global _lib
    if _lib is None:
        if not _LIB.exists():
            build_library()
        _lib = ctypes.CDLL(str(_LIB))
    return _lib


def _ptr(array: np.ndarray):
    return array.ctypes.data_as(ctypes.c_void_p)


def _suffix(dtype) -> str:
    return {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}[np.dtype(dtype)]


def eps(dtype) -> float:
    return float(np.finfo(dtype).eps)


def grid_dims(lengths, radius: float, dtype) -> list[int]:
    """``_neighbor_list.py:115-122``: bins per axis, host side, in Python float64."""
    e = eps(dtype)
    return [max(1, math.floor(float(L) / (radius + 4 * e * float(L)))) for L in lengths]


def all_pairs_path(dims) -> bool:
    """``_neighbor_list.py:131-140``: the all-pairs path iff the first grid has <= 3 bins per axis."""
    return max(dims) <= 3


def wrap(x: np.ndarray, cell: np.ndarray):
    """``space.py:22-38`` in the dtype of ``x``: (wrapped, crossings int64)."""
    x = np.ascontiguousarray(x)
    cell = np.ascontiguousarray(cell, dtype=x.dtype)
    out = np.empty_like(x)
    crossings = np.empty(x.shape, dtype=np.int64)
    getattr(lib(), f"orc_wrap_{_suffix(x.dtype)}")(
        _ptr(x), ctypes.c_int64(x.shape[0]), _ptr(cell), _ptr(out), _ptr(crossings)
    )
    return out, crossings


def capacity_for(count: int, multiplier: float) -> int:
    """``_neighbor_list.py:259``: C = max(8, ceil(m * count))."""
    return max(8, math.ceil(multiplier * count))


class Build:
    """One build's outputs, named as the reference's buffers."""

    def __init__(self, **fields):
        self.__dict__.update(fields)


def build(
    x_wrapped: np.ndarray,
    cell: np.ndarray,
    radius: float,
    *,
    dims=None,
    widths=None,
    all_pairs: bool | None = None,
    capacity: int | None = None,
    capacity_multiplier: float = 1.25,
) -> Build:
    """Count and fill a list for wrapped positions [N, 3] in the reference's column order.

    ``dims``/``widths`` default to the first-frame rule (host float64, then cast); pass the
    update's device values to follow an update. ``capacity`` defaults to the allocate rule.
    Returns pairs [2, C] int32 with padding (k mod N, k mod N), pair_total, counts, offsets,
    sorted_ids (grid path) and the overflow status, as the reference's count, fill, total and
    pad kernels leave them (``pairs.cuh:114-157``).
    """
    x = np.ascontiguousarray(x_wrapped)
    dtype = x.dtype
    cell = np.ascontiguousarray(cell, dtype=dtype)
    n = x.shape[0]
    suf = _suffix(dtype)
    L = lib()
    if dims is None:
        dims = grid_dims(cell.tolist(), radius, dtype)
    dims = np.ascontiguousarray(dims, dtype=np.int32)
    if all_pairs is None:
        all_pairs = all_pairs_path(dims)
    counts = np.zeros(n, dtype=np.int32)
    sorted_ids = atom_bin = bin_offsets = sp = None
    if all_pairs:
        getattr(L, f"orc_all_pairs_{suf}")(
            _ptr(x), _ptr(cell), ctypes.c_double(radius), ctypes.c_int32(n),
            _ptr(counts), None, ctypes.c_int64(0), None,
        )
    else:
        if widths is None:
            widths = np.array([float(Lx) / int(d) for Lx, d in zip(cell.tolist(), dims)], dtype=dtype)
        widths = np.ascontiguousarray(widths, dtype=dtype)
        bins = int(np.prod(dims))
        atom_bin = np.empty(n, dtype=np.int32)
        sorted_ids = np.empty(n, dtype=np.int32)
        bin_offsets = np.empty(bins + 1, dtype=np.int32)
        sp = np.empty_like(x)
        getattr(L, f"orc_grid_{suf}")(
            _ptr(x), ctypes.c_int32(n), _ptr(dims), _ptr(widths),
            _ptr(atom_bin), _ptr(sorted_ids), _ptr(bin_offsets), _ptr(sp),
        )
        getattr(L, f"orc_count_grid_{suf}")(
            _ptr(sp), _ptr(sorted_ids), _ptr(atom_bin), _ptr(bin_offsets), _ptr(dims),
            _ptr(cell), ctypes.c_double(radius), ctypes.c_int32(n), _ptr(counts),
        )
    offsets = np.zeros(n, dtype=np.int32)
    np.cumsum(counts[:-1], out=offsets[1:]) if n > 1 else None
    total = int(counts.astype(np.int64).sum())
    if capacity is None:
        capacity = capacity_for(total, capacity_multiplier)
    pairs = np.full((2, capacity), -1, dtype=np.int32)
    if all_pairs:
        getattr(L, f"orc_all_pairs_{suf}")(
            _ptr(x), _ptr(cell), ctypes.c_double(radius), ctypes.c_int32(n),
            None, _ptr(offsets), ctypes.c_int64(capacity), _ptr(pairs),
        )
    else:
        getattr(L, f"orc_fill_grid_{suf}")(
            _ptr(sp), _ptr(sorted_ids), _ptr(atom_bin), _ptr(bin_offsets), _ptr(dims),
            _ptr(cell), ctypes.c_double(radius), ctypes.c_int32(n), _ptr(offsets),
            ctypes.c_int64(capacity), _ptr(pairs),
        )
    status = 0
    if total > capacity:
        status = OVERFLOW
    else:
        k = np.arange(total, capacity, dtype=np.int64)
        pairs[0, total:] = (k % n).astype(np.int32)
        pairs[1, total:] = (k % n).astype(np.int32)
    return Build(
        pairs=pairs, pair_total=total, counts=counts, offsets=offsets, sorted_ids=sorted_ids,
        atom_bin=atom_bin, bin_offsets=bin_offsets, sorted_positions=sp, dims=dims,
        capacity=capacity, status=status, all_pairs=all_pairs,
    )


def prepare_update(cell, reference_cell, cutoff, radius, limits, status=0):
    """``update.cuh:27-75`` in the dtype of ``cell``: (rebuild, status, dims, widths, reuse)."""
    cell = np.ascontiguousarray(cell)
    dtype = cell.dtype
    ref = np.ascontiguousarray(reference_cell, dtype=dtype)
    limits = np.ascontiguousarray(limits, dtype=np.int32)
    dims = np.zeros(3, dtype=np.int32)
    widths = np.zeros(3, dtype=dtype)
    reuse = np.zeros(4, dtype=dtype)
    rebuild = np.zeros(1, dtype=np.int32)
    st = np.array([status, 0], dtype=np.int32)
    getattr(lib(), f"orc_prepare_update_{_suffix(dtype)}")(
        _ptr(cell), _ptr(ref), ctypes.c_double(cutoff), ctypes.c_double(radius), _ptr(limits),
        _ptr(dims), _ptr(widths), _ptr(reuse), _ptr(rebuild), _ptr(st),
    )
    return int(rebuild[0]), int(st[0]), dims, widths, reuse


def check_reuse(x, x0, cell, reuse, rebuild: int, status: int = 0):
    """``update.cuh:79-116``: (rebuild, status) after the per-atom reuse check."""
    x = np.ascontiguousarray(x)
    dtype = x.dtype
    x0 = np.ascontiguousarray(x0, dtype=dtype)
    cell = np.ascontiguousarray(cell, dtype=dtype)
    reuse = np.ascontiguousarray(reuse, dtype=dtype)
    rb = np.array([rebuild], dtype=np.int32)
    st = np.array([status, 0], dtype=np.int32)
    getattr(lib(), f"orc_check_reuse_{_suffix(dtype)}")(
        _ptr(x), _ptr(x0), _ptr(cell), _ptr(reuse), ctypes.c_int32(x.shape[0]), _ptr(rb), _ptr(st)
    )
    return int(rb[0]), int(st[0])


def update_decision(x_wrapped, cell, reference_positions, reference_cell, cutoff, radius, limits):
    """The update's rebuild decision: prepare, then the reuse check (``entry_points.cu:466-488``).

    Returns (rebuild, status, dims, widths).
    """
    rebuild, status, dims, widths, reuse = prepare_update(
        cell, reference_cell, cutoff, radius, limits
    )
    if status == 0:
        rebuild, status = check_reuse(x_wrapped, reference_positions, cell, reuse, rebuild, status)
    return rebuild, status, dims, widths


def minimum_image_np(center, other, length):
    """``traversal.cuh:31-38`` elementwise in the operands' dtype."""
    dtype = center.dtype
    raw = center - other
    half = dtype.type(0.5) * length
    neg_half = dtype.type(-0.5) * length
    return np.where(
        raw > half, (center - length) - other, np.where(raw < neg_half, center - (other - length), raw)
    )


def brute_pairs(x_wrapped: np.ndarray, cell: np.ndarray, radius: float, block: int = 2048):
    """Every unordered pair (i < j) with min-image r^2 < fl(R)^2, as a sorted [P, 2] int64 array.

    Independent of the C restatement: a plain O(N^2) numpy search in the dtype of ``x``.
    """
    x = np.ascontiguousarray(x_wrapped)
    dtype = x.dtype
    cell = np.asarray(cell, dtype=dtype)
    r = dtype.type(radius)
    r2 = r * r
    n = x.shape[0]
    out = []
    for start in range(0, n, block):
        stop = min(n, start + block)
        c = x[start:stop, None, :]
        d = minimum_image_np(c, x[None, :, :], cell)
        d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        i, j = np.nonzero(d2 < r2)
        i = i + start
        keep = j > i
        out.append(np.stack([i[keep], j[keep]], axis=1))
    pairs = np.concatenate(out) if out else np.zeros((0, 2), dtype=np.int64)
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return pairs[order].astype(np.int64)


def pair_set(pairs: np.ndarray, total: int) -> np.ndarray:
    """The first ``total`` columns of a [2, C] list as sorted unordered (min, max) rows."""
    first = pairs[0, :total].astype(np.int64)
    second = pairs[1, :total].astype(np.int64)
    rows = np.stack([np.minimum(first, second), np.maximum(first, second)], axis=1)
    order = np.lexsort((rows[:, 1], rows[:, 0]))
    return rows[order]


def jittered_lattice(per_axis: int, length: float, jitter: float, seed: int, dtype=np.float64):
    """A simple-cubic lattice of per_axis^3 sites in a cubic box, jittered uniform +-jitter.

    Generated in fp64 from ``numpy.random.default_rng(seed)`` and cast last; ids run x fastest.
    """
    a = length / per_axis
    g = np.arange(per_axis, dtype=np.float64)
    z, y, xx = np.meshgrid(g, g, g, indexing="ij")
    sites = np.stack([xx.ravel(), y.ravel(), z.ravel()], axis=1) * a + 0.5 * a
    rng = np.random.default_rng(seed)
    sites = sites + rng.uniform(-jitter, jitter, size=sites.shape)
    return sites.astype(dtype)
