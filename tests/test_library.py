"""The C-ABI library loads and exports every symbol its header declares; the torch ops carry
the reference's schemas. No GPU compute here."""

import ctypes

import pytest
import torch

from nabladyn_amd import _native
from nabladyn_amd.neighbors import NeighborList, _ops


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    declared = _native.exported_symbols()
    assert len(declared) >= 8
    missing = [name for name in declared if not hasattr(lib, name)]
    assert missing == []


def test_scan_temp_bytes_and_argument_errors():
    out = ctypes.c_int64()
    _native.call("scan_temp_bytes", ctypes.c_int64(1000), ctypes.byref(out))
    assert out.value >= 256 + 4
    with pytest.raises(_native.NativeError, match="fit int32"):
        _native.call("scan_temp_bytes", ctypes.c_int64(0), ctypes.byref(out))
    buffers = _native.NeighborBuffers()
    with pytest.raises(_native.NativeError, match="N must satisfy"):
        _native.call("count_first_frame", 0, None, None, ctypes.c_double(1.0),
                     ctypes.byref(buffers), None)


def test_md_batch_argument_errors():
    from nabladyn_amd.md import _batch

    batch = _batch._Batch()
    batch.dtype = _native.F32
    members = (_batch._Member * 1)()
    with pytest.raises(_native.NativeError, match="1 to 256"):
        _native.call("md_batch_prepare", ctypes.byref(batch), members, 0, None)
    with pytest.raises(_native.NativeError, match="structure is null"):
        _native.call("md_batch_prepare", ctypes.byref(batch), members, 1, None)
    batch.dtype = 7
    with pytest.raises(_native.NativeError, match="dtype"):
        _native.call("md_batch_prepare", ctypes.byref(batch), members, 1, None)
    with pytest.raises(_native.NativeError, match="not prepared"):
        _native.call("md_batch_step", ctypes.byref(_batch._Batch()), None)
    assert ctypes.sizeof(_batch._Batch) == 8 * 2 + 4 * 6 + 8 * 3 + 8 * 3 * _batch._KERNELS


def test_op_schemas_match_the_reference():
    ops = _ops.ops()
    reference = {
        "scan_temp_bytes": "nabladyn_neighbors::scan_temp_bytes(Tensor reference, int items) -> int",
    }
    schema = str(ops.scan_temp_bytes.default._schema)
    assert schema == reference["scan_temp_bytes"]
    update = str(ops.update.default._schema)
    for name in ("Tensor(o!) reference_positions", "Tensor(t!) bits", "float cutoff, float radius",
                 "Tensor limits", "Tensor(q!) reuse"):
        assert name in update
    assert ops.scan_temp_bytes(torch.empty(1), 100) > 0


def test_neighbor_list_validation_on_host():
    with pytest.raises(ValueError):
        NeighborList(0.0)
    with pytest.raises(ValueError):
        NeighborList(1.0, skin=-0.1)
    with pytest.raises(ValueError):
        NeighborList(1.0, capacity_multiplier=0.5)
    with pytest.raises(ValueError):
        NeighborList(1.0, inner_cutoffs=(1.5,))
    nl = NeighborList(1.0, skin=0.1)
    assert nl.radius == pytest.approx(1.1)
    with pytest.raises(TypeError):
        nl.allocate(torch.zeros(4, 3), torch.ones(3) * 3)
    with pytest.raises(RuntimeError, match="allocate"):
        nl.tensors()


def test_noalloc_entry_point_refuses_what_its_scratch_cannot_hold():
    """ndn_pair_efw_noalloc's limit: a table whose padded knots outgrow the static scratch is
    refused with NDN_ERR_ARGUMENT before anything runs (no device memory is touched: the
    pointers below are never read)."""
    pot = _native.Potential(kind=_native.TABLE, num_types=6, atom_types=ctypes.c_void_p(8),
                            table=ctypes.c_void_p(8), start=ctypes.c_void_p(8),
                            spacing=ctypes.c_void_p(8), knots=8000, cutoff=1.2)
    with pytest.raises(_native.NativeError, match="scratch"):
        _native.call("pair_efw_noalloc", _native.F64, ctypes.c_void_p(8), ctypes.c_void_p(8),
                     ctypes.c_int64(1000), ctypes.c_void_p(8), ctypes.c_int64(5000),
                     ctypes.byref(pot), None, None, None, None)


def test_pair_entry_points_refuse_a_sum():
    """``next`` (a sum of two models) is for the MD entry points; the pair entry points refuse
    it with NDN_ERR_ARGUMENT before anything runs, and ndn_md_workspace sizes both tables."""
    second = _native.Potential(kind=_native.LJ, num_types=1, sigma=ctypes.c_void_p(8),
                               epsilon=ctypes.c_void_p(8), cutoff=0.8)
    pot = _native.Potential(kind=_native.TABLE, num_types=1, table=ctypes.c_void_p(8),
                            start=ctypes.c_void_p(8), spacing=ctypes.c_void_p(8), knots=200,
                            cutoff=1.0, next=ctypes.cast(ctypes.pointer(second), ctypes.c_void_p))
    size = ctypes.c_int64()
    with pytest.raises(_native.NativeError, match="sum of potentials"):
        _native.call("pair_workspace_bytes", _native.F32, ctypes.byref(pot), ctypes.c_int64(64),
                     ctypes.byref(size))
    slots, alone, both = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _native.call("md_workspace", _native.F32, ctypes.byref(pot), ctypes.c_int64(100),
                 ctypes.byref(slots), ctypes.byref(both))
    pot.next = None
    _native.call("md_workspace", _native.F32, ctypes.byref(pot), ctypes.c_int64(100),
                 ctypes.byref(slots), ctypes.byref(alone))
    # the table's 202 padded knots and spacing (fp32, 256-byte multiples), then an LJ row
    assert alone.value == 1024 and both.value == 1024 + 256


def test_force_launch_refuses_too_few_partial_slots():
    """The host guard of the MD force launch: one fp64 partial slot per block (a block per brick,
    rounded up to a multiple of 8) must fit the system's partials, else NDN_ERR_ARGUMENT before
    any launch (the pointers below are never read)."""
    from nabladyn_amd.md._simulation import MDList, MDSystem

    fake = ctypes.c_void_p(256)
    system = MDSystem(num_atoms=1000, positions=fake, velocities=fake, forces=fake, images=fake,
                      masses=fake, inverse_masses=fake, cell=fake, energy=fake, virial=fake,
                      thermostat=fake, chain_masses=fake, kinetic=fake, partials=fake,
                      partial_slots=8, coefficients=fake, coefficient_bytes=4096)
    md_list = MDList(group=4, stride=64, scan_temp_bytes=1 << 20, halo_capacity=100, bricks=75,
                     **{name: fake for name in (
                         "bin_counts", "bin_offsets", "bin_of", "unordered", "order", "counts",
                         "index", "ids", "positions4", "scratch", "scan_temp", "status",
                         "prepare_dims", "prepare_widths", "prepare_total", "halo_slots",
                         "halo_runs", "halo_counts", "grid_widths")})
    md_list.dims[:] = (20, 20, 20)
    md_list.brick[:] = (4, 4, 8)  # 5 x 5 x 3 = 75 bricks: 80 blocks
    pot = _native.Potential(kind=_native.LJ, num_types=1, sigma=fake, epsilon=fake, cutoff=1.0)
    with pytest.raises(_native.NativeError, match="too few partial slots"):
        _native.call("md_force", _native.F32, ctypes.byref(system), ctypes.byref(pot),
                     ctypes.byref(md_list), 0, 1, ctypes.c_double(0.004), None)
    system.partial_slots = 80
    system.coefficient_bytes = 16
    with pytest.raises(_native.NativeError, match="coefficient buffer"):
        _native.call("md_force", _native.F32, ctypes.byref(system), ctypes.byref(pot),
                     ctypes.byref(md_list), 0, 1, ctypes.c_double(0.004), None)
