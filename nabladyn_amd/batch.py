"""Many independent systems of uneven size in one set of launches (``ndn_batch_efw``).

``Batch`` takes systems — (positions [n, 3], cell [3], model), each model a single-type
LennardJones or Tabulated potential of this package — and evaluates energy, forces and virial of
all of them with one binning pass, one scan and one force launch. ``schedule`` is "auto" (the
autotuner times the four variants on this batch once and keeps the fastest), or one of
("thread", 1), ("thread", 2), ("wave", 1), ("wave", 2): a lane or a wave per atom, and sub-bins
at least the cutoff or half the cutoff wide.

``plain_efw`` is the per-system path the batch is measured against: for each system, a
``NeighborList`` (zero skin) allocated on it and ``quantity.energy_force_virial`` — the
reference's own way of evaluating one system (``examples/lj_evaluate.py``).
"""

import math
import time
from dataclasses import dataclass

import numpy as np
import torch
from torch import Tensor

from . import _native, quantity
from .neighbors import NeighborList
from .potential import LennardJones, Tabulated

SCHEDULES = (("thread", 1), ("thread", 2), ("wave", 1), ("wave", 2))
_TUNED: dict[tuple, tuple[str, int]] = {}


class NativeBatch(_native.ctypes.Structure):
    """``ndn_batch``."""

    _vp, _i64, _i32 = _native.ctypes.c_void_p, _native.ctypes.c_int64, _native.ctypes.c_int32
    _fields_ = [
        ("num_systems", _i64), ("num_atoms", _i64), ("positions", _vp), ("system_of", _vp),
        ("atom_start", _vp), ("bin_start", _vp), ("dims", _vp), ("widths", _vp), ("cells", _vp),
        ("cutoffs", _vp), ("cutoffs_squared", _vp), ("kinds", _vp), ("lj_rows", _vp), ("table_of", _vp), ("tables", _vp),
        ("table_start", _vp), ("table_spacing", _vp), ("knots", _i64),
        ("total_bins", _i64), ("span", _i32), ("reserved", _i32), ("bin_of", _vp),
        ("bin_counts", _vp), ("bin_offsets", _vp), ("unordered", _vp), ("order", _vp),
        ("wrapped4", _vp), ("sorted4", _vp), ("scan_temp", _vp), ("scan_temp_bytes", _i64),
        ("status", _vp), ("forces", _vp), ("atom_terms", _vp), ("energy", _vp), ("virial", _vp),
    ]


def lj_rows(model: LennardJones, dtype) -> Tensor:
    """sigma^2, 4 eps, 24 eps and the energy at the cutoff, in the model's dtype, as the
    reference's ``_pair_energy(cutoff**2, ...)`` gives it (``_lennard_jones/_potential.py:113-120``)."""
    sigma, eps = model.sigma.detach()[0].to(dtype), model.epsilon.detach()[0].to(dtype)
    s2 = sigma * sigma
    base = s2 / torch.tensor(model.cutoff**2, dtype=dtype, device=sigma.device)
    p6 = base * base * base
    return torch.stack([s2, 4 * eps, 24 * eps, 4 * eps * (p6 * p6 - p6)])


def padded_table(model: Tabulated, dtype) -> tuple[Tensor, Tensor, Tensor]:
    """[K + 2] knots of row 0 with the virtual end knots and the last knot 0, its start and
    spacing (``_tabulated/_potential.py:79-82``)."""
    v = model.table.detach()[0].to(dtype).clone()
    v[-1] = 0
    before = 3 * v[0] - 3 * v[1] + v[2]
    after = 3 * v[-1] - 3 * v[-2] + v[-3]
    padded = torch.cat([before[None], v, after[None]])
    return padded, model.start[0].to(dtype), model.spacing[0].to(dtype)


def sub_bin_dims(lengths, cutoff: float, span: int, dtype) -> list[int]:
    eps = torch.finfo(dtype).eps
    width = cutoff / span * (1.0 + 1e-6)
    return [max(1, math.floor(float(L) / (width + 4 * eps * float(L)))) for L in lengths]


@dataclass
class _Layout:
    span: int
    native: NativeBatch
    keep: dict


class Batch:
    """Energy, forces and virial of many systems per launch; see the module docstring."""

    def __init__(self, systems, *, schedule="auto"):
        if not systems:
            raise ValueError("a batch needs at least one system")
        positions, cells, models = zip(*systems)
        self.dtype = positions[0].dtype
        self.device = positions[0].device
        _native.require_gpu(positions[0], "positions")
        for x, cell, model in systems:
            if x.dtype != self.dtype or x.device != self.device or x.ndim != 2 or x.shape[1] != 3:
                raise ValueError("every system's positions must be [n, 3] of one dtype and device")
            if model.num_atoms != x.shape[0]:
                raise ValueError("a model must match its system's atom count")
            if getattr(model, "_num_types", 1) != 1:
                raise ValueError("batched systems take one atom type each")
            if not isinstance(model, (LennardJones, Tabulated)):
                raise TypeError("batched systems take LennardJones or Tabulated models")
            lengths = [float(v) for v in cell.tolist()]
            if not all(math.isfinite(v) and v > 2 * model.cutoff for v in lengths):
                raise ValueError("every cell length must be finite and above twice the cutoff")
        self.sizes = [int(x.shape[0]) for x in positions]
        self.models = list(models)
        self.cells = torch.stack([c.to(self.dtype) for c in cells]).to(self.device).contiguous()
        self.positions = torch.cat([x.detach() for x in positions]).contiguous()
        self.num_atoms = self.positions.shape[0]
        if self.num_atoms >= 2**31:
            raise ValueError("a batch holds fewer than 2**31 atoms")
        counts = torch.tensor(self.sizes, dtype=torch.int64)
        self.atom_start = torch.cat([torch.zeros(1, dtype=torch.int64), counts.cumsum(0)]).to(self.device)
        self.system_of = torch.repeat_interleave(
            torch.arange(len(systems), dtype=torch.int32), counts).to(self.device)
        self._lengths = self.cells.cpu().tolist()
        self._prepare_potentials()
        n, s = self.num_atoms, len(systems)
        self.forces = torch.zeros((n, 3), dtype=self.dtype, device=self.device)
        self.atom_terms = torch.zeros((n, 7), dtype=self.dtype, device=self.device)
        self.energy = torch.zeros(s, dtype=self.dtype, device=self.device)
        self.virial = torch.zeros((s, 3, 3), dtype=self.dtype, device=self.device)
        self._layouts: dict[int, _Layout] = {}
        self.schedule = schedule
        if schedule != "auto" and tuple(schedule) not in SCHEDULES:
            raise ValueError(f"schedule must be 'auto' or one of {SCHEDULES}")

    def _prepare_potentials(self) -> None:
        dt, dev = self.dtype, self.device
        kinds, rows, table_of, tables, starts, spacings, cutoffs = [], [], [], [], [], [], []
        knots = 0
        for model in self.models:
            cutoffs.append(model.cutoff)
            if isinstance(model, LennardJones):
                kinds.append(_native.LJ)
                rows.append(lj_rows(model, dt))
                table_of.append(-1)
            else:
                kinds.append(_native.TABLE)
                rows.append(torch.zeros(4, dtype=dt, device=dev))
                padded, start, spacing = padded_table(model, dt)
                if knots and padded.numel() != knots + 2:
                    raise ValueError("every table of a batch must have the same number of knots")
                knots = padded.numel() - 2
                table_of.append(len(tables))
                tables.append(padded)
                starts.append(start)
                spacings.append(spacing)
        ints = {"dtype": torch.int32, "device": dev}
        self.kinds = torch.tensor(kinds, **ints)
        self.lj = torch.stack(rows).contiguous()
        self.table_of = torch.tensor(table_of, **ints)
        self.tables = torch.stack(tables).contiguous() if tables else torch.zeros((1, 4), dtype=dt, device=dev)
        self.table_start = torch.stack(starts) if starts else torch.zeros(1, dtype=dt, device=dev)
        self.table_spacing = torch.stack(spacings) if spacings else torch.ones(1, dtype=dt, device=dev)
        self.knots = knots
        self.cutoffs = torch.tensor(cutoffs, dtype=dt, device=dev)
        # squared in double, rounded once to the dtype: the reference's `squared < cutoff**2`
        self.cutoffs_squared = torch.tensor([c * c for c in cutoffs], dtype=dt, device=dev)

    def _layout(self, span: int) -> _Layout:
        if span in self._layouts:
            return self._layouts[span]
        dims = [sub_bin_dims(L, m.cutoff, span, self.dtype) for L, m in zip(self._lengths, self.models)]
        bins = [math.prod(d) for d in dims]
        total = sum(bins)
        if total >= 2**31 - 1:
            raise ValueError("too many bins in the batch")
        dev, n = self.device, self.num_atoms
        ints = {"dtype": torch.int32, "device": dev}
        bin_start = torch.tensor(np.concatenate([[0], np.cumsum(bins)]), **ints)
        widths = torch.tensor([[L[a] / d[a] for a in range(3)] for L, d in zip(self._lengths, dims)],
                              dtype=self.dtype, device=dev)
        keep = {
            "dims": torch.tensor(dims, **ints), "widths": widths, "bin_start": bin_start,
            "bin_of": torch.zeros(n, **ints), "bin_counts": torch.zeros(total + 1, **ints),
            "bin_offsets": torch.zeros(total + 1, **ints), "unordered": torch.zeros(n, **ints),
            "order": torch.zeros(n, **ints),
            "wrapped4": torch.zeros((n, 4), dtype=self.dtype, device=dev),
            "sorted4": torch.zeros((n, 4), dtype=self.dtype, device=dev),
            "status": torch.zeros(2, **ints),
        }
        scan = _native.lib().ndn_batch_scan_bytes
        scan.restype = _native.ctypes.c_int64
        keep["scan_temp"] = torch.zeros(int(scan(_native.ctypes.c_int64(total))), dtype=torch.uint8,
                                        device=dev)
        p = _native.ptr
        native = NativeBatch(
            num_systems=len(self.sizes), num_atoms=n, positions=p(self.positions),
            system_of=p(self.system_of), atom_start=p(self.atom_start), bin_start=p(bin_start),
            dims=p(keep["dims"]), widths=p(widths), cells=p(self.cells), cutoffs=p(self.cutoffs),
            cutoffs_squared=p(self.cutoffs_squared),
            kinds=p(self.kinds), lj_rows=p(self.lj), table_of=p(self.table_of),
            tables=p(self.tables) if self.knots else None, table_start=p(self.table_start),
            table_spacing=p(self.table_spacing), knots=self.knots, total_bins=total,
            span=span, bin_of=p(keep["bin_of"]), bin_counts=p(keep["bin_counts"]),
            bin_offsets=p(keep["bin_offsets"]), unordered=p(keep["unordered"]),
            order=p(keep["order"]), wrapped4=p(keep["wrapped4"]), sorted4=p(keep["sorted4"]),
            scan_temp=p(keep["scan_temp"]), scan_temp_bytes=keep["scan_temp"].numel(),
            status=p(keep["status"]), forces=p(self.forces), atom_terms=p(self.atom_terms),
            energy=p(self.energy), virial=p(self.virial),
        )
        layout = _Layout(span, native, keep)
        self._layouts[span] = layout
        return layout

    def signature(self) -> tuple:
        return (str(self.dtype), tuple(self.sizes), tuple(type(m).__name__ for m in self.models))

    def launch(self, schedule: tuple[str, int]) -> None:
        """One evaluation with this schedule, on the current stream, no sync."""
        layout = self._layout(schedule[1])
        _native.call("batch_efw", _native.dtype_tag(self.dtype), _native.ctypes.byref(layout.native),
                     0 if schedule[0] == "thread" else 1, _native.stream(self.device))

    def tune(self, repeats: int = 3) -> tuple[str, int]:
        """The autotuner: times each schedule on this batch (HIP events) and keeps the fastest.
        Cached per batch signature. Syncs."""
        key = self.signature()
        if key in _TUNED:
            return _TUNED[key]
        timings = {}
        for schedule in SCHEDULES:
            self.launch(schedule)
            start = torch.cuda.Event(enable_timing=True)
            end = torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(repeats):
                self.launch(schedule)
            end.record()
            torch.cuda.synchronize(self.device)
            timings[schedule] = start.elapsed_time(end) / repeats
        best = min(timings, key=timings.get)
        self.timings = timings
        _TUNED[key] = best
        return best

    def evaluate(self) -> tuple[Tensor, Tensor, Tensor]:
        """(energy [S], forces [N, 3] in the systems' order, virial [S, 3, 3])."""
        schedule = self.tune() if self.schedule == "auto" else tuple(self.schedule)
        self.chosen = schedule
        self.launch(schedule)
        return self.energy, self.forces, self.virial

    def split(self, tensor: Tensor) -> list[Tensor]:
        return list(torch.split(tensor, self.sizes))

    def set_positions(self, positions: list[Tensor]) -> None:
        """New frames for the same systems (same sizes), copied in place."""
        torch.cat([x.detach().to(self.dtype) for x in positions], out=self.positions)


def make_systems(raw, device, dtype=torch.float32, table=None):
    """(positions, cell, model) for each dict of ``workloads.ragged_systems``: argon LJ
    (sigma 0.34 nm, epsilon 0.996 kJ/mol, cutoff 1.0 nm) or the 200-knot table on
    [0.3, 1.2] nm (``workloads.lj_table``), one atom type."""
    from . import potential, workloads

    table = workloads.lj_table() if table is None else table
    out = []
    for system in raw:
        n = system["positions"].shape[0]
        x = torch.tensor(system["positions"], dtype=dtype, device=device)
        cell = torch.tensor(system["cell"], dtype=dtype, device=device)
        types = torch.zeros(n, dtype=torch.int64, device=device)
        if system["kind"] == "lj":
            model = potential.lennard_jones(
                atom_types=types,
                sigma=torch.tensor([workloads.ARGON_SIGMA], dtype=dtype, device=device),
                epsilon=torch.tensor([workloads.ARGON_EPSILON], dtype=dtype, device=device),
                cutoff=1.0)
        else:
            model = potential.tabulated(
                atom_types=types, table=torch.tensor(table[None], dtype=dtype, device=device),
                start=torch.tensor([0.3], dtype=dtype, device=device), cutoff=1.2)
        out.append((x, cell, model))
    return out


def plain_efw(systems):
    """The per-system path: a zero-skin NeighborList and quantity.energy_force_virial per system
    (the reference's way of evaluating one frame). Returns [(E, F, W)]."""
    out = []
    for x, cell, model in systems:
        nl = NeighborList(model.cutoff, skin=0.0)
        nl.allocate(x, cell)
        out.append(quantity.energy_force_virial(model)(x, cell, nl.tensors()))
    return out


def time_call(fn, repeats: int = 3) -> float:
    """Wall seconds per call of fn (synchronized)."""
    fn()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(repeats):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - start) / repeats


__all__ = ("Batch", "plain_efw", "SCHEDULES")
