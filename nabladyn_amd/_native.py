"""The gfx950 library: ctypes bindings of its C ABI (include/nabladyn_amd.h).

The library is ``_lib/libnabladyn_amd.so``, built in-tree by ``nabladyn_amd.build``. Loading
fails loudly when it is missing: there is no other implementation of these operations.
Every call passes raw device pointers and torch's current HIP stream, so the kernels run in
stream order with the surrounding torch work and can be captured in a HIP graph.
"""

import ctypes
import os
import threading
from pathlib import Path

import torch

_LIBRARY = Path(__file__).resolve().parent / "_lib" / "libnabladyn_amd.so"
_LOCK = threading.Lock()
_lib = None

F32, F64 = 0, 1
LJ, TABLE = 0, 1
STATUS_INVALID_CELL, STATUS_NON_FINITE, STATUS_OVERFLOW, STATUS_FULL_OVERFLOW = 1, 2, 8, 16

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64


class NeighborBuffers(ctypes.Structure):
    """``ndn_neighbor_buffers``."""

    _fields_ = [
        ("dims", _vp), ("widths", _vp), ("atom_bin", _vp), ("bin_counts", _vp),
        ("bin_offsets", _vp), ("unordered_ids", _vp), ("sorted_ids", _vp),
        ("sorted_positions", _vp), ("scan_temp", _vp), ("pair_counts", _vp),
        ("pair_offsets", _vp), ("pair_total", _vp), ("status", _vp), ("pairs", _vp),
        ("bits", _vp), ("num_atoms", _i64), ("bin_slots", _i64), ("scan_temp_bytes", _i64),
        ("capacity", _i64), ("bits_words", _i64),
        ("full_counts", _vp), ("full_slices", _vp), ("full_index", _vp), ("full_capacity", _i64),
    ]


class Potential(ctypes.Structure):
    """``ndn_potential``."""

    _fields_ = [
        ("kind", ctypes.c_int), ("num_types", ctypes.c_int), ("atom_types", _vp),
        ("sigma", _vp), ("epsilon", _vp), ("table", _vp), ("start", _vp), ("spacing", _vp),
        ("knots", _i64), ("cutoff", ctypes.c_double), ("next", _vp),
    ]


class NativeError(RuntimeError):
    """A C-ABI call rejected its arguments or failed to launch."""


def library_path() -> Path:
    """The library to load: ``NABLADYN_AMD_LIBRARY`` if set (another build of the same ABI),
    else the in-tree build."""
    override = os.environ.get("NABLADYN_AMD_LIBRARY")
    return Path(override) if override else _LIBRARY


def lib() -> ctypes.CDLL:
    """Load the library once per process; raise ``RuntimeError`` if it was not built."""
    global _lib
    if _lib is None:
        with _LOCK:
            if _lib is None:
                path = library_path()
                if not path.exists():
                    raise RuntimeError(
                        f"{path} is missing: build it with `python -m nabladyn_amd.build`"
                    )
                handle = ctypes.CDLL(str(path))
                handle.ndn_last_error.restype = ctypes.c_char_p
                _lib = handle
    return _lib


def call(name: str, *args) -> None:
    """Call ``ndn_<name>``; raise ``NativeError`` with the library's message on failure."""
    check(name, getattr(lib(), "ndn_" + name)(*args))


def check(name: str, code: int) -> None:
    """Raise ``NativeError`` with the library's message when ``ndn_<name>`` returned ``code``
    other than 0 (for callers that hold the function itself)."""
    if code != 0:
        message = lib().ndn_last_error().decode()
        raise NativeError(f"ndn_{name}: {message}")


def ptr(tensor: torch.Tensor | None):
    """The device address of a tensor (None for None or an empty tensor)."""
    if tensor is None or tensor.numel() == 0:
        return None
    return ctypes.c_void_p(tensor.data_ptr())


def dtype_tag(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return F32
    if dtype == torch.float64:
        return F64
    raise TypeError(f"expected float32 or float64, got {dtype}")


def stream(device: torch.device | None = None):
    """torch's current HIP stream on the device, as a raw handle."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_gpu(tensor: torch.Tensor, name: str = "tensor") -> None:
    if not tensor.is_cuda:
        raise TypeError(f"{name} must be on a GPU (a ROCm 'cuda' device)")


def exported_symbols() -> list[str]:
    """The C-ABI names the header declares, as the library must export them."""
    header = Path(__file__).resolve().parents[1] / "include" / "nabladyn_amd.h"
    import re

    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(ndn_\w+)\s*\(", header.read_text(), re.M)))
