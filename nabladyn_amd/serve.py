"""Continuous-batching service of energy/force/virial and short-MD requests on one GPU.

Requests arrive over time (``Server.submit``). Each scheduling round the server:

1. admits waiting requests that have arrived, in arrival order, while the resident set's device
   memory stays within the HBM budget (a request's footprint is estimated before admission and
   measured once it is resident);
2. launches one batched evaluation (``batch.Batch``: one binning pass, one scan, one force
   launch) for every energy/force request admitted since the last round, on the evaluation
   stream;
3. advances every resident MD request by a chunk of steps: the resident simulations are
   grouped by ``md.batch_key`` and kind of potential into ``md.SimulationBatch`` es (one launch
   per kernel of the step for the whole group), each group on its own MD stream, so evaluations
   and MD segments overlap on the device and many small systems cost the launches of one;
4. waits for the round's work, completes what finished and frees its memory.

A request that arrives while others are on the GPU joins at the next round, as memory frees up.
Latency is completion minus arrival, on the server's clock.

Every result is bitwise the same alone and in the mix: the evaluation batch reduces each system
in a fixed order, independent of the other systems, under a fixed schedule; in an MD batch each
member gets the blocks its own step launches and a fixed force schedule.

There is no reference counterpart; the baseline is ``serve_plain``, each request alone on the
plain path (a ``NeighborList`` plus ``quantity.energy_force_virial``, or ``md.nvt`` and ``run``).
"""

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch
from torch import Tensor

from . import md, potential, workloads
from .batch import Batch, plain_efw
from .potential import LennardJones, Tabulated

# Device bytes per atom and per request, an upper estimate used for admission before a request's
# real footprint is known: an evaluation holds its positions, the batch's sorted copies, forces
# and per-atom terms; a simulation holds its frame, the permutation scratch and its full list.
EFW_BYTES_PER_ATOM = 256
MD_BYTES_PER_ATOM = 1536
REQUEST_BYTES = 1 << 20


@dataclass
class Request:
    """One request: an evaluation ("efw") of E, F, W, or an NVT segment ("md") of ``steps``."""

    kind: str
    positions: Tensor
    cell: Tensor
    model: object
    velocities: Tensor | None = None
    masses: Tensor | None = None
    steps: int = 0
    temperature: float = 120.0
    thermostat_timescale: float = 0.1
    timestep: float = 0.004
    skin: float = 0.2
    arrival: float = 0.0
    # filled by the server
    id: int = -1
    admitted: float | None = None
    finished: float | None = None
    footprint: int = 0
    result: dict = field(default_factory=dict)

    @property
    def atoms(self) -> int:
        return int(self.positions.shape[0])

    @property
    def atom_steps(self) -> int:
        return self.atoms * (self.steps if self.kind == "md" else 1)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class _Running:
    """A resident MD request: its simulation and the steps left."""

    def __init__(self, request: Request, simulation):
        self.request, self.simulation = request, simulation
        self.left = request.steps


class Server:
    """See the module docstring. ``hbm_budget`` is in bytes; ``md_chunk`` steps at most per MD
    request per round; ``md_streams`` streams for the MD batches (all at the default priority);
    ``md_group`` the force schedule of every MD request (fixed, so results do not depend on
    timing)."""

    def __init__(self, device, *, hbm_budget: int = 16 << 30, md_chunk: int = 100,
                 md_streams: int = 8, efw_schedule=("wave", 2), md_group: int = 4):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the server runs on a GPU device")
        if hbm_budget <= 0 or md_chunk <= 0 or md_streams <= 0:
            raise ValueError("hbm_budget, md_chunk and md_streams must be positive")
        self.hbm_budget = int(hbm_budget)
        self.md_chunk = int(md_chunk)
        self.efw_schedule = tuple(efw_schedule)
        self.md_group = int(md_group)
        self.efw_stream = torch.cuda.Stream(self.device)
        self.md_streams = [torch.cuda.Stream(self.device) for _ in range(md_streams)]
        self.waiting: list[Request] = []
        self.running: list[_Running] = []
        self.completed: list[Request] = []
        self.resident_bytes = 0
        self.peak_resident_bytes = 0
        self.rounds = 0
        self._next_id = 0
        self._batches: dict[tuple, md.SimulationBatch] = {}

    # ---- requests --------------------------------------------------------------------------

    def estimate(self, request: Request) -> int:
        per_atom = MD_BYTES_PER_ATOM if request.kind == "md" else EFW_BYTES_PER_ATOM
        return REQUEST_BYTES + per_atom * request.atoms

    def submit(self, request: Request) -> int:
        """Queues a request; returns its id. Raises ValueError for a request that could never be
        admitted (its footprint alone exceeds the budget) or is malformed."""
        if request.kind not in ("efw", "md"):
            raise ValueError("a request's kind is 'efw' or 'md'")
        if not isinstance(request.model, (LennardJones, Tabulated)):
            raise TypeError("requests take LennardJones or Tabulated models")
        if request.positions.device != self.device:
            raise ValueError("a request's tensors must be on the server's device")
        if request.kind == "md" and (request.steps <= 0 or request.velocities is None
                                     or request.masses is None):
            raise ValueError("an md request needs steps > 0, velocities and masses")
        if self.estimate(request) > self.hbm_budget:
            raise ValueError("the request alone exceeds the HBM budget")
        request.id = self._next_id
        self._next_id += 1
        self.waiting.append(request)
        return request.id

    # ---- one round -------------------------------------------------------------------------

    def _admit(self, now: float) -> tuple[list[Request], list[_Running]]:
        evaluations, started = [], []
        while self.waiting and self.waiting[0].arrival <= now:
            request = self.waiting[0]
            need = self.estimate(request)
            if self.resident_bytes + need > self.hbm_budget:
                break  # arrival order: the head waits for memory, the rest wait behind it
            self.waiting.pop(0)
            request.admitted = now
            request.footprint = need
            self.resident_bytes += need
            if request.kind == "efw":
                evaluations.append(request)
            else:
                before = torch.cuda.memory_allocated(self.device)
                simulation = md.Simulation(
                    request.model, request.positions, request.velocities, request.masses,
                    request.cell, timestep=request.timestep, skin=request.skin,
                    capacity_multiplier=1.25, images=None, step=0,
                    temperature=request.temperature,
                    thermostat_timescale=request.thermostat_timescale,
                    force_group=self.md_group, graphs=False)
                measured = torch.cuda.memory_allocated(self.device) - before
                if measured > need:  # account the real footprint
                    self.resident_bytes += measured - need
                    request.footprint = measured
                started.append(_Running(request, simulation))
            self.peak_resident_bytes = max(self.peak_resident_bytes, self.resident_bytes)
        return evaluations, started

    def round(self, now: float) -> list[Request]:
        """Admit, launch, wait, complete. Returns the requests completed in this round."""
        evaluations, started = self._admit(now)
        self.running.extend(started)
        if not evaluations and not self.running:
            return []
        self.rounds += 1
        producer = torch.cuda.current_stream(self.device)
        batch = None
        if evaluations:
            self.efw_stream.wait_stream(producer)
            with torch.cuda.stream(self.efw_stream):
                batch = Batch([(r.positions, r.cell, r.model) for r in evaluations],
                              schedule=self.efw_schedule)
                energy, forces, virial = batch.evaluate()
                parts = batch.split(forces)
                for k, request in enumerate(evaluations):
                    request.result = {"energy": energy[k].clone(), "forces": parts[k].clone(),
                                      "virial": virial[k].clone()}
        # each group's batch is kept while its members stay and remade when they change
        batches = {}
        for k, group in enumerate(self._md_groups()):
            stream = self.md_streams[k % len(self.md_streams)]
            stream.wait_stream(producer)
            steps = min(self.md_chunk, min(run.left for run in group))
            key = tuple(id(run.simulation) for run in group)
            with torch.cuda.stream(stream):
                stepper = self._batches.get(key)
                if stepper is None:  # its preparation syncs the stream
                    stepper = md.SimulationBatch([run.simulation for run in group])
                batches[key] = stepper
                stepper.advance(steps)
            for run in group:
                run.left -= steps
        self._batches = batches
        torch.cuda.synchronize(self.device)
        done = []
        if batch is not None:
            del batch
            done.extend(evaluations)
        still = []
        for run in self.running:
            if run.left > 0:
                still.append(run)
                continue
            run.simulation.check()
            state = run.simulation.state()
            request = run.request
            request.result = {"positions": state.positions, "velocities": state.velocities,
                              "forces": state.forces, "images": state.images}
            done.append(request)
        self.running = still
        finished = self.clock()
        for request in done:
            request.finished = finished
            self.resident_bytes -= request.footprint
        self.completed.extend(done)
        return done

    def _md_groups(self) -> list[list[_Running]]:
        """The resident MD requests by batch key and kind of potential, in admission order, at
        most 256 per group: one batch per kind, each on its own stream, so the kinds' steps
        overlap on the device (one batch of mixed kinds runs its force launches one after the
        other, with the LDS of the largest halo in its shared launches)."""
        groups: dict[tuple, list[_Running]] = {}
        for run in self.running:
            pot = run.simulation._pot
            key = md.batch_key(run.simulation) + (pot.kind, pot.num_types == 1)
            groups.setdefault(key, []).append(run)
        out = []
        for members in groups.values():
            out.extend(members[k:k + 256] for k in range(0, len(members), 256))
        return out

    # ---- the loop --------------------------------------------------------------------------

    def clock(self) -> float:
        return time.perf_counter() - self._start

    def run(self, requests: list[Request]) -> dict:
        """Serves `requests` (sorted by arrival, in seconds from now) in real time and returns
        the statistics of ``summarize``."""
        pending = sorted(requests, key=lambda r: r.arrival)
        torch.cuda.synchronize(self.device)
        self._start = time.perf_counter()
        while pending or self.waiting or self.running:
            now = self.clock()
            while pending and pending[0].arrival <= now:
                self.submit(pending.pop(0))
            if not self.running and not any(r.arrival <= now for r in self.waiting):
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                continue
            self.round(now)
        return summarize(self.completed, self.clock(), rounds=self.rounds,
                         peak_resident_bytes=self.peak_resident_bytes)


def summarize(completed: list[Request], wall: float, **extra) -> dict:
    latencies = np.array([r.latency for r in completed]) if completed else np.zeros(1)
    atom_steps = sum(r.atom_steps for r in completed)
    out = {
        "requests": len(completed),
        "wall_s": wall,
        "p50_latency_ms": float(np.percentile(latencies, 50) * 1e3),
        "p99_latency_ms": float(np.percentile(latencies, 99) * 1e3),
        "atom_steps_per_s": atom_steps / wall if wall > 0 else math.nan,
        "requests_per_s": len(completed) / wall if wall > 0 else math.nan,
    }
    out.update(extra)
    return out


def serve_plain(requests: list[Request], *, md_group: int = 4) -> dict:
    """The per-system baseline: each request alone, in arrival order, on the plain path — a
    zero-skin ``NeighborList`` and ``quantity.energy_force_virial`` for an evaluation,
    ``md.nvt`` and ``run`` for a segment — with the same clock and statistics as ``Server``."""
    pending = sorted(requests, key=lambda r: r.arrival)
    device = pending[0].positions.device
    torch.cuda.synchronize(device)
    start = time.perf_counter()
    for request in pending:
        now = time.perf_counter() - start
        if request.arrival > now:
            time.sleep(request.arrival - now)
        if request.kind == "efw":
            energy, forces, virial = plain_efw([(request.positions, request.cell, request.model)])[0]
            request.result = {"energy": energy, "forces": forces, "virial": virial}
        else:
            simulation = md.Simulation(
                request.model, request.positions, request.velocities, request.masses,
                request.cell, timestep=request.timestep, skin=request.skin,
                capacity_multiplier=1.25, images=None, step=0, temperature=request.temperature,
                thermostat_timescale=request.thermostat_timescale, force_group=md_group)
            simulation.run(request.steps)
            state = simulation.state()
            request.result = {"positions": state.positions, "velocities": state.velocities,
                              "forces": state.forces, "images": state.images}
        torch.cuda.synchronize(device)
        request.finished = time.perf_counter() - start
    return summarize(pending, time.perf_counter() - start)


def make_requests(specs: list[dict], device, dtype=torch.float32, *, min_length: float = 2.45,
                  table=None) -> list[Request]:
    """Requests from ``workloads.request_stream`` specs: a jittered lattice of the spec's size at
    21 nm^-3 in a box with aspect ratios in [1, 2] (every edge at least ``min_length``, above
    twice the table's 1.2 nm cutoff), an LJ (argon) or a 200-knot table potential, and for MD
    Maxwell-Boltzmann velocities at 120 K from the spec's seed."""
    table = workloads.lj_table() if table is None else table
    out = []
    for spec in specs:
        rng = np.random.default_rng(spec["seed"])
        n = spec["atoms"]
        ratios = np.array([1.0, rng.uniform(1, 2), rng.uniform(1, 2)])
        lengths = ratios * (n / workloads.DENSITY / ratios.prod()) ** (1 / 3)
        if lengths.min() < min_length:
            lengths = lengths * (min_length / lengths.min())
        per = np.maximum(1, np.round(lengths / (lengths.prod() / n) ** (1 / 3))).astype(int)
        while per.prod() < n:
            per[np.argmin(per / lengths)] += 1
        x = workloads.lattice(per, lengths, 0.02, rng, n=n)
        positions = torch.tensor(x, dtype=dtype, device=device)
        cell = torch.tensor(lengths, dtype=dtype, device=device)
        types = torch.zeros(n, dtype=torch.int64, device=device)
        if spec["potential"] == "lj":
            model = potential.lennard_jones(
                atom_types=types, sigma=torch.tensor([workloads.ARGON_SIGMA], dtype=dtype, device=device),
                epsilon=torch.tensor([workloads.ARGON_EPSILON], dtype=dtype, device=device), cutoff=1.0)
        else:
            model = potential.tabulated(
                atom_types=types, table=torch.tensor(table[None], dtype=dtype, device=device),
                start=torch.tensor([0.3], dtype=dtype, device=device), cutoff=1.2)
        request = Request(spec["kind"], positions, cell, model, arrival=spec["arrival"],
                          steps=spec["steps"])
        if spec["kind"] == "md":
            masses = torch.full((n,), workloads.ARGON_MASS, dtype=dtype, device=device)
            request.masses = masses
            request.velocities = md.maxwell_boltzmann(masses, 120.0, spec["seed"]).to(dtype)
        out.append(request)
    return out


__all__ = ("Request", "Server", "make_requests", "serve_plain", "summarize")
