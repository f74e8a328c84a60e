"""Benchmark: the 2,097,152-atom liquid-argon LJ NVT step on one MI355X.

Workload: a jittered 128^3 simple-cubic lattice of argon at 21 nm^-3 (cubic box 46.3949 nm),
fp32, LJ sigma 0.34 nm, epsilon 0.996 kJ/mol, cutoff 1.0 nm, skin 0.2 nm, Nose-Hoover chain of 3
at 120 K with a 0.1 ps timescale, 4 fs steps; each step one HIP graph replay with the rebuild
decided on the device. Synthetic, random-seeded inputs (no datasets).

Prints one JSON line: atom-steps/s over K timed steps (after W untimed ones), the force
kernel's roofline (algorithmic bytes / its event-timed average duration, against 8 TB/s), the
batched engine's frames/s against the plain per-system path (256 ragged systems), batched MD's
atom-steps/s against the plain per-system step graphs (the serving stream's 14 NVT systems), a
sum of two potentials' E/F/W on the 2M frame with and without an inner list, the reference's own
benchmark (LJ NVT at 512 .. 32,768 atoms, ns/day beside its published RTX 5090 figures), one
update of a 64-frame force-matching fit of a table, the serving loop's p50/p99 latency and
throughput against the plain per-request path, and a CPU baseline (the oracle's restatement of
the reference path on this host's cores).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-batching]
                    [--no-serving]
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12  # B/s, MI355X
MIN_TIMED_STEPS = 1000
MIN_WARMUP_STEPS = 200
PER_AXIS = 128
CUTOFF, SKIN = 1.0, 0.2
METRIC = "atom-steps/s (2M-atom LJ NVT) + batched frames/s vs plain; p99 req latency"


def log(*parts):
    print(*parts, file=sys.stderr, flush=True)


def build_simulation(torch, dtype=None, group=None):
    from nabladyn_amd import md, potential, workloads

    dtype = dtype or torch.float32
    device = torch.device("cuda", 0)
    x, cell = workloads.argon_lattice(PER_AXIS, seed=0)
    n = x.shape[0]
    masses = torch.full((n,), workloads.ARGON_MASS, dtype=dtype)
    velocities = md.maxwell_boltzmann(masses, 120.0, 1)
    model = potential.lennard_jones(
        atom_types=torch.zeros(n, dtype=torch.int64, device=device),
        sigma=torch.tensor([workloads.ARGON_SIGMA], dtype=dtype, device=device),
        epsilon=torch.tensor([workloads.ARGON_EPSILON], dtype=dtype, device=device),
        cutoff=CUTOFF,
    )
    sim = md.nvt(
        model,
        torch.tensor(x, dtype=dtype, device=device),
        velocities.to(device),
        masses.to(device),
        torch.tensor(cell, dtype=dtype, device=device),
        timestep=0.004,
        temperature=120.0,
        thermostat_timescale=0.1,
        skin=SKIN,
    )
    if group is not None:
        # a fixed schedule (the counters' runs: a profiler's overhead can sway the autotuner)
        sim = md.Simulation(
            model, torch.tensor(x, dtype=dtype, device=device), velocities.to(device),
            masses.to(device), torch.tensor(cell, dtype=dtype, device=device), timestep=0.004,
            skin=SKIN, capacity_multiplier=1.25, images=None, step=0, temperature=120.0,
            thermostat_timescale=0.1, force_group=group)
    return sim, x, cell


def time_force_kernel(torch, sim, launches=50):
    """Average duration of the step's force kernel, timed with HIP events on its stream."""
    from nabladyn_amd import _native

    def launch():
        _native.call(
            "md_force", _native.dtype_tag(sim._dtype), _native.ctypes.byref(sim._system),
            _native.ctypes.byref(sim._pot), _native.ctypes.byref(sim._neighbors.native), 0, 1,
            _native.ctypes.c_double(sim._timestep), _native.stream(sim._device),
        )

    for _ in range(3):
        launch()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(launches):
        launch()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1e3 / launches


def bench_batching(torch, count=256, repeats=10):
    """256 ragged systems (log-uniform 64..32768 atoms, aspect ratios 1..2, half LJ, half a
    200-knot table; edges stretched to >= 2.45 nm where 21 nm^-3 would make them shorter than
    twice the table's cutoff), fp32 E/F/W: the batch engine (one set of launches, schedule
    autotuned) against the plain per-system path (NeighborList + energy_force_virial each)."""
    from nabladyn_amd import batch, workloads

    device = torch.device("cuda", 0)
    raw = workloads.ragged_systems(count, seed=2, min_length=2.45)
    systems = batch.make_systems(raw, device)
    b = batch.Batch(systems, schedule="auto")
    b.evaluate()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(repeats):
        b.launch(b.chosen)
    end.record()
    torch.cuda.synchronize()
    t_batch = start.elapsed_time(end) / 1e3 / repeats
    batch.plain_efw(systems[:4])  # warm the plain path's launches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.plain_efw(systems)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    atoms = sum(int(x.shape[0]) for x, _, _ in systems)
    return {
        "systems": count,
        "atoms": atoms,
        "schedule": list(b.chosen),
        "batched_frames_per_s": count / t_batch,
        "plain_frames_per_s": count / t_plain,
        "speedup_vs_plain": t_plain / t_batch,
        "batched_ms_per_set": t_batch * 1e3,
        "plain_ms_per_set": t_plain * 1e3,
    }


def bench_serving(torch, count=64, qps=40.0):
    """A Poisson stream of `count` requests at `qps` (70% E/F/W, 30% 500-step NVT segments,
    256..262144 atoms log-uniform, LJ or 200-knot table, fp32): the continuous-batching server's
    p50/p99 latency and throughput, and the throughput of the same requests all queued at once
    on the server and on the plain per-request path. Each all-queued run follows one untimed run
    of the same requests on the same path, so the allocator's pool already holds what they need
    (the first run in a process grows it), as it does for a server in service."""
    from nabladyn_amd import serve, workloads

    device = torch.device("cuda", 0)
    specs = workloads.request_stream(count, qps, seed=4)
    live = serve.Server(device).run(serve.make_requests(specs, device))
    queued = [dict(spec, arrival=0.0) for spec in specs]
    serve.Server(device).run(serve.make_requests(queued, device))
    capacity = serve.Server(device).run(serve.make_requests(queued, device))
    serve.serve_plain(serve.make_requests(queued, device))
    plain = serve.serve_plain(serve.make_requests(queued, device))
    return {
        "requests": count,
        "qps": qps,
        "md_requests": sum(spec["kind"] == "md" for spec in specs),
        "p50_latency_ms": live["p50_latency_ms"],
        "p99_latency_ms": live["p99_latency_ms"],
        "atom_steps_per_s": live["atom_steps_per_s"],
        "capacity_atom_steps_per_s": capacity["atom_steps_per_s"],
        "plain_atom_steps_per_s": plain["atom_steps_per_s"],
        "speedup_vs_plain": capacity["atom_steps_per_s"] / plain["atom_steps_per_s"],
        "plain_p50_latency_ms": plain["p50_latency_ms"],
        "plain_p99_latency_ms": plain["p99_latency_ms"],
    }


def bench_md_batch(torch, steps=200, warmup=20, plain=True):
    """The serving stream's MD requests (its 14 NVT systems of 262..205,875 atoms, 736,295 in
    all, LJ or a 200-knot table, fp32), stepped in one ``md.SimulationBatch`` per kind of
    potential, each on its own stream, against the plain path: each system alone, its own
    captured step graphs replayed (``Simulation.run``), one after another. Both from the same
    frames after ``warmup`` steps; atom-steps/s over ``steps``."""
    from nabladyn_amd import md, serve, workloads

    device = torch.device("cuda", 0)
    specs = [s for s in workloads.request_stream(64, 40.0, seed=4) if s["kind"] == "md"]

    def simulations(graphs):
        out = []
        for r in serve.make_requests([dict(s, arrival=0.0) for s in specs], device):
            out.append(md.Simulation(
                r.model, r.positions, r.velocities, r.masses, r.cell, timestep=r.timestep,
                skin=r.skin, capacity_multiplier=1.25, images=None, step=0,
                temperature=r.temperature, thermostat_timescale=r.thermostat_timescale,
                force_group=4, graphs=graphs))
        return out

    sims = simulations(graphs=False)
    atoms = sum(s._num_atoms for s in sims)
    kinds = {}
    for s in sims:
        kinds.setdefault((s._pot.kind, s._pot.num_types == 1), []).append(s)
    batches = [md.SimulationBatch(group) for group in kinds.values()]
    streams = [torch.cuda.Stream(device) for _ in batches]

    def advance(count):
        producer = torch.cuda.current_stream(device)
        for batch, stream in zip(batches, streams):
            stream.wait_stream(producer)
            with torch.cuda.stream(stream):
                batch.advance(count)
        for stream in streams:
            producer.wait_stream(stream)

    advance(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    advance(steps)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    for s in sims:
        s.check()
    out = {
        "systems": len(sims),
        "atoms": atoms,
        "batches": len(batches),
        "steps": steps,
        "batched_ms_per_step": t_batch / steps * 1e3,
        "batched_atom_steps_per_s": atoms * steps / t_batch,
    }
    if plain:
        alone = simulations(graphs=True)
        for s in alone:
            s.run(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in alone:
            s.run(steps)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        out.update(plain_ms_per_step=t_plain / steps * 1e3,
                   plain_atom_steps_per_s=atoms * steps / t_plain,
                   speedup_vs_plain=t_plain / t_batch)
    return out


def bench_sum(torch, x, cell, repeats=5):
    """``potential.sum`` on the 2,097,152-atom argon frame (fp32): LJ (argon, rc 0.9 nm) plus a
    200-knot table (rc 1.2 nm, the LJ's shape), E/F/W through ``quantity.energy_force_virial``
    on a zero-skin list with an inner list at 0.9 nm and on the 1.2 nm list alone. Returns the
    mean time of each over ``repeats`` evaluations (HIP events) and the rate against the
    algorithmic bytes: each part's 8 B per listed pair within its cutoff plus 24 B per atom (x
    read, F written), as for an E/F/W evaluation."""
    from nabladyn_amd import potential, quantity, workloads
    from nabladyn_amd.neighbors import NeighborList

    device = torch.device("cuda", 0)
    n = x.shape[0]
    xd = torch.tensor(x, dtype=torch.float32, device=device)
    cd = torch.tensor(cell, dtype=torch.float32, device=device)
    types = torch.zeros(n, dtype=torch.int64, device=device)
    lj = potential.lennard_jones(
        atom_types=types, sigma=torch.tensor([workloads.ARGON_SIGMA], device=device),
        epsilon=torch.tensor([workloads.ARGON_EPSILON], device=device), cutoff=0.9)
    table = potential.tabulated(
        atom_types=types,
        table=torch.tensor(workloads.lj_table()[None], dtype=torch.float32, device=device),
        start=torch.tensor([0.3], device=device), cutoff=1.2)
    model = potential.sum(lj, table)
    efw = quantity.energy_force_virial(model)
    out = {"atoms": n, "parts": "LJ rc 0.9 nm + 200-knot table rc 1.2 nm", "dtype": "fp32"}
    for label, inner in (("inner_list", (0.9,)), ("one_list", ())):
        nl = NeighborList(1.2, skin=0.0, inner_cutoffs=inner)
        nl.allocate(xd, cd)
        pairs = nl.tensors()
        efw(xd, cd, pairs)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record()
        for _ in range(repeats):
            efw(xd, cd, pairs)
        end.record()
        torch.cuda.synchronize()
        seconds = start.elapsed_time(end) / 1e3 / repeats
        out[f"{label}_ms"] = seconds * 1e3
        if inner:
            listed = [int(pairs.inner[0].index.shape[1]), int(pairs.index.shape[1])]
            near = nl._inner[0].num_pairs
            far = nl.num_pairs
            algorithmic = 8 * (near + far) + 2 * 24 * n
            out.update(pairs_within=[near, far], algorithmic_bytes=algorithmic,
                       achieved_GBps=algorithmic / seconds / 1e9,
                       frac=algorithmic / seconds / HBM_PEAK, columns=listed)
        del nl, pairs
    torch.cuda.empty_cache()
    return out


def bench_fitting(torch, repeats=5, cpu_frames=2):
    """Force matching of a table: 64 frames of 4,096 beads (8 nm^-3, L = 8 nm), a one-type
    200-knot table on [0.3, 1.2] nm (last knot 0), targets from LJ(0.47 nm, 3.5 kJ/mol) at
    rc 1.2 nm (forces plus N(0, 5) noise; E and W exact), fp32, loss E + F + W. Times one
    update's batch gradient over all 64 frames (HIP events, mean of ``repeats``) through
    ``training`` (a zero-skin pair stack per frame and full rows built from it once; one
    ndn_pair_efw_full_frames forward and one ndn_pair_parameter_vjp_frames backward for all
    the frames) and through ``fit.TableFrames`` (the batched engine, all frames in one launch,
    and ndn_batch_table_vjp), and the oracle's autograd
    double backward on ``cpu_frames`` frames on the host. Reports the loss in the reference's
    scaled form and as plain MSE, and the two GPU routes' gradient distance."""
    import numpy as np

    from nabladyn_amd import fit, potential, training, workloads
    from nabladyn_amd.training import _fit

    device = torch.device("cuda", 0)
    frames, knots, start, cutoff = 64, 200, 0.3, 1.2
    xs, cells, hidden = workloads.cg_frames(frames=frames, beads=4096)
    beads = xs.shape[1]
    positions64 = torch.tensor(xs, dtype=torch.float64, device=device)
    cells64 = torch.tensor(cells, dtype=torch.float64, device=device)
    types = torch.zeros(beads, dtype=torch.int64, device=device)
    truth = potential.lennard_jones(
        atom_types=types, sigma=torch.tensor([hidden["sigma"]], dtype=torch.float64, device=device),
        epsilon=torch.tensor([hidden["epsilon"]], dtype=torch.float64, device=device),
        cutoff=cutoff)
    stack64 = training.pair_stack(positions64, cells64, cutoff)
    with torch.no_grad():
        labels = training.predict(truth, {}, positions64, cells64, (stack64,))
    noise = hidden["rng"].normal(0.0, hidden["force_noise"], (frames, beads, 3))
    labels["forces"] = labels["forces"] + torch.tensor(noise, device=device)
    labels = {k: v.float() for k, v in labels.items()}
    r = start + (cutoff - start) / (knots - 1) * np.arange(knots)
    s6 = (0.45 / r) ** 6
    table = np.clip(4 * 3.0 * (s6 * s6 - s6), -50, 50)
    table[-1] = 0.0
    model = potential.tabulated(
        atom_types=types, table=torch.tensor(table[None], dtype=torch.float32, device=device),
        start=torch.tensor([start], dtype=torch.float32, device=device), cutoff=cutoff,
        trainable=("table",))
    weights = {"energy": 1.0, "forces": 1.0, "virial": 1.0}
    data = training.Dataset(xs.astype(np.float32), cells.astype(np.float32),
                            **{k: v.cpu() for k, v in labels.items()})
    t0 = time.perf_counter()
    prepared = _fit._prepare(model, data, training.CHANNELS)
    torch.cuda.synchronize()
    prepare_s = time.perf_counter() - t0
    scales = training.scales(**prepared.labels)
    trained = {"table": model.table}
    everything = torch.arange(frames)

    def timed(fn):
        fn()
        begin, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        begin.record()
        for _ in range(repeats):
            result = fn()
        end.record()
        torch.cuda.synchronize()
        return begin.elapsed_time(end) / repeats, result

    per_frame_ms, (gradient, scaled, _) = timed(lambda: _fit.batch_gradient(
        model, trained, prepared, everything, frames, weights, scales))
    batched = fit.TableFrames(model, prepared.positions, prepared.cells)

    def batched_gradient():
        t = model.table.detach().clone().requires_grad_(True)
        value, _ = fit.loss(batched.predict(t), prepared.labels, weights, scales)
        (g,) = torch.autograd.grad(value, t)
        return g

    batched_ms, batched_grad = timed(batched_gradient)
    with torch.no_grad():
        predicted = training.predict(model, {}, prepared.positions, prepared.cells,
                                     prepared.stacks)
    plain = {k: float((predicted[k].double() - prepared.labels[k].double()).square().mean())
             for k in training.CHANNELS}
    g = gradient["table"].double()
    distance = float((g - batched_grad.double()).abs().max() / g.abs().max())
    # the oracle's double backward (the reference's route) on a few frames, on the host
    from oracle import neighbors as on
    from oracle import potentials as op
    from oracle import training as ot

    sample = []
    for f in range(cpu_frames):
        xw = on.wrap(xs[f].astype(np.float32).astype(np.float64), cells[f])[0]
        built = on.build(xw, cells[f], cutoff)
        index = torch.tensor(built.pairs[:, :built.pair_total]).long()
        sample.append((torch.tensor(xw, dtype=torch.float32),
                       torch.tensor(cells[f], dtype=torch.float32), index))
    spacing = torch.tensor([(cutoff - start) / (knots - 1)])
    cpu_types = torch.zeros(beads, dtype=torch.int64)

    def energy(p, x, strain, cell, index):
        return op.t_table_energy(x, strain, cell, index, cpu_types, p[0],
                                 torch.tensor([start]), spacing, cutoff)

    cpu_labels = {k: v[:cpu_frames].cpu() for k, v in prepared.labels.items()}
    t0 = time.perf_counter()
    ot.loss_and_gradient(energy, [torch.tensor(table[None], dtype=torch.float32)], sample,
                         cpu_labels, weights, scales)
    cpu_ms = (time.perf_counter() - t0) * 1e3 / cpu_frames
    return {
        "workload": "64 frames x 4,096 beads, one-type 200-knot table, E+F+W, fp32",
        "pairs_columns": int(prepared.stacks[0].shape[2]),
        "prepare_s": prepare_s,
        "update_ms_training": per_frame_ms,
        "frames_per_s_training": frames / per_frame_ms * 1e3,
        "update_ms_batched_engine": batched_ms,
        "frames_per_s_batched_engine": frames / batched_ms * 1e3,
        "gradient_distance_between_routes": distance,
        "loss_scaled": float(scaled),
        "loss_plain_mse": plain,
        "cpu_oracle_ms_per_frame": cpu_ms,
        "cpu_cores": torch.get_num_threads(),
    }


# The reference's published LJ NVT figures (1x RTX 5090, fp32, dt 4 fs, skin 0.1 nm, rc 1.0 nm,
# Nose-Hoover chain of 3 at 120 K; README.md:594-603): atoms -> ns/day.
REFERENCE_NS_PER_DAY = {512: 9288.0, 4096: 5507.0, 13824: 3632.0, 32768: 2210.0}


def bench_published_sizes(torch, steps=2000, warmup=200):
    """The reference's own benchmark configuration at its four published sizes: liquid-argon LJ
    NVT, fp32, jittered simple-cubic lattices of 8^3 .. 32^3 atoms at 21 nm^-3, rc 1.0 nm,
    skin 0.1 nm, 4 fs, chain of 3 at 120 K, each step one graph replay (``md.nvt`` + ``run``).
    ns/day and atom-steps/s over ``steps`` steps after ``warmup``, beside the reference's figure
    (another GPU: context, not a like-for-like target)."""
    from nabladyn_amd import md, potential, workloads

    device = torch.device("cuda", 0)
    out = []
    for n, published in REFERENCE_NS_PER_DAY.items():
        per_axis = round(n ** (1 / 3))
        x, cell = workloads.argon_lattice(per_axis, seed=0)
        masses = torch.full((n,), workloads.ARGON_MASS)
        velocities = md.maxwell_boltzmann(masses, 120.0, 1)
        model = potential.lennard_jones(
            atom_types=torch.zeros(n, dtype=torch.int64, device=device),
            sigma=torch.tensor([workloads.ARGON_SIGMA], device=device),
            epsilon=torch.tensor([workloads.ARGON_EPSILON], device=device), cutoff=1.0)
        sim = md.nvt(model, torch.tensor(x, dtype=torch.float32, device=device),
                     velocities.float().to(device), masses.to(device),
                     torch.tensor(cell, dtype=torch.float32, device=device), timestep=0.004,
                     temperature=120.0, thermostat_timescale=0.1, skin=0.1)
        sim.run(warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run(steps)
        torch.cuda.synchronize()
        seconds = (time.perf_counter() - t0) / steps
        ns_per_day = 0.004e-3 * 86400 / seconds
        out.append({"atoms": n, "us_per_step": seconds * 1e6, "ns_per_day": ns_per_day,
                    "atom_steps_per_s": n / seconds,
                    "reference_ns_per_day_rtx5090": published,
                    "ratio_to_reference": ns_per_day / published})
        del sim
    return out


def cpu_baseline(x, cell, steps_per_rebuild):
    """The oracle (the reference's eager torch force path + C cell list) on this host's cores:
    one forced rebuild and one force evaluation plus the Verlet/chain update of the same
    2,097,152-atom frame; atom-steps/s with the rebuild amortized over the GPU run's measured
    steps per rebuild."""
    import numpy as np
    import torch

    from oracle import neighbors as on
    from oracle import potentials as op

    threads = torch.get_num_threads()
    n = x.shape[0]
    x32, L32 = x.astype(np.float32), cell.astype(np.float32)
    t0 = time.perf_counter()
    xw, _ = on.wrap(x32, L32)
    b = on.build(xw, L32, CUTOFF + SKIN)
    t_build = time.perf_counter() - t0
    index = torch.as_tensor(b.pairs).long()
    xt, Lt = torch.tensor(xw), torch.tensor(L32)
    types = torch.zeros(n, dtype=torch.int64)
    sigma, eps = torch.tensor([0.34]), torch.tensor([0.996])

    def energy(pos, strain, c, idx):
        return op.t_lj_energy(pos, strain, c, idx, types, sigma, eps, CUTOFF)

    grad = torch.func.grad_and_value(energy, argnums=0)
    v = torch.zeros_like(xt)
    t0 = time.perf_counter()
    gx, e = grad(xt, torch.zeros(3, 3), Lt, index)
    forces = -gx
    v = v + 0.002 * forces / 39.948
    xt2 = torch.remainder(xt + 0.004 * v, Lt)
    kinetic = 0.5 * (39.948 * v.square()).sum()
    t_force = time.perf_counter() - t0
    # without a measured cadence the rebuild is charged to the one step
    per_rebuild = steps_per_rebuild if steps_per_rebuild else 1.0
    step = t_force + t_build / per_rebuild
    del index, gx, forces, xt2, kinetic
    return {
        "value": n / step,
        "unit": "atom-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (
            f"one 2,097,152-atom fp32 LJ NVT step: a forced cell-list rebuild ({t_build:.2f} s, "
            f"1 thread, C restatement) amortized over {per_rebuild:.1f} steps per rebuild (the GPU "
            f"run's measured cadence), "
            f"plus the eager torch-autograd force and update ({t_force:.2f} s, {threads} threads)"
        ),
    }


def cpu_host():
    """The host the CPU figures ran on: logical CPUs, the CPU model, torch's threads."""
    import torch

    model = None
    try:
        with open("/proc/cpuinfo") as info:
            model = next((line.split(":", 1)[1].strip() for line in info
                          if line.startswith("model name")), None)
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "torch_threads": torch.get_num_threads()}


def cpu_small_configs(nve_steps=200, systems=16):
    """The planned CPU figures beside the 2M one, on bounded samples: the 512-atom argon LJ
    frame (jittered 8^3 lattice, L = 2.90 nm, fp32, skin 0.1 nm) stepped NVE ``nve_steps`` times
    by the oracle's eager integrator (atom-steps/s), and the per-system loop of E/F/W over the
    first ``systems`` of the 256 ragged systems (the oracle's C cell list at zero skin, then its
    autograd E/F/W), in systems/s and atom-evaluations/s of that sample."""
    import numpy as np
    import torch

    from nabladyn_amd import md, workloads
    from oracle import md as omd
    from oracle import neighbors as on
    from oracle import potentials as op

    dtype = torch.float32
    x = torch.tensor(on.jittered_lattice(8, 2.90, 0.02, seed=0), dtype=dtype)
    masses = torch.full((512,), workloads.ARGON_MASS, dtype=dtype)
    v = md.maxwell_boltzmann(masses, 120.0, 1).to(dtype)
    model = omd.LJModel(np.zeros(512, np.int64), [workloads.ARGON_SIGMA],
                        [workloads.ARGON_EPSILON], CUTOFF, dtype)
    integrator = omd.Integrator(model, x, v, masses, torch.full((3,), 2.90, dtype=dtype),
                                dt=0.004, skin=0.1)
    t0 = time.perf_counter()
    for _ in range(nve_steps):
        integrator.step()
    t_nve = time.perf_counter() - t0
    raw = workloads.ragged_systems(256, seed=2, min_length=2.45)[:systems]
    table = torch.tensor(workloads.lj_table()[None], dtype=dtype)
    start, table_cutoff = torch.tensor([0.3]), 1.2
    spacing = (table_cutoff - start) / (table.shape[1] - 1)
    atoms = 0
    t0 = time.perf_counter()
    for system in raw:
        xs = system["positions"].astype(np.float32)
        cell = system["cell"].astype(np.float32)
        xw, _ = on.wrap(xs, cell)
        n = xw.shape[0]
        atoms += n
        types = torch.zeros(n, dtype=torch.int64)
        cutoff = CUTOFF if system["kind"] == "lj" else table_cutoff
        built = on.build(xw, cell, cutoff)
        index = torch.as_tensor(built.pairs[:, :built.pair_total]).long()
        if system["kind"] == "lj":
            args = (types, torch.tensor([workloads.ARGON_SIGMA]),
                    torch.tensor([workloads.ARGON_EPSILON]), CUTOFF)
            op.torch_efw(op.t_lj_energy, torch.tensor(xw), torch.tensor(cell), index, *args)
        else:
            op.torch_efw(op.t_table_energy, torch.tensor(xw), torch.tensor(cell), index, types,
                         table, start, spacing, table_cutoff)
    t_loop = time.perf_counter() - t0
    return {
        "nve_512_atom_steps_per_s": 512 * nve_steps / t_nve,
        "nve_512_sample": f"{nve_steps} fp32 NVE steps of the 512-atom argon frame",
        "ragged_systems_per_s": systems / t_loop,
        "ragged_atom_evals_per_s": atoms / t_loop,
        "ragged_sample": f"E/F/W of the first {systems} of the 256 ragged systems ({atoms} atoms), "
                         "fp32",
    }


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=2000)
    parser.add_argument("--warmup", type=int, default=200)
    parser.add_argument("--no-cpu-baseline", action="store_true")
    parser.add_argument("--no-batching", action="store_true")
    parser.add_argument(
        "--kernel", choices=("force", "copy", "md-batch", "sum", "published", "fitting"),
        help="only one measurement, for counter collection and traces: 'force' builds the "
             "2M-atom system (4 lanes per row, the autotuner's choice for it), runs 100 steps, "
             "then launches its force kernel 23 times (the last 23 dispatches of that kernel); 'copy' copies 1 GiB on the device 20 times (the "
             "scale of the traffic counters); 'md-batch' steps the serving stream's 14 MD "
             "systems in batches (20 + 200 steps) and prints that leg alone; 'sum' evaluates "
             "potential.sum on the 2M frame and prints that leg alone; 'published' runs the "
             "reference's published sizes and prints that leg alone; 'fitting' times one "
             "update of the 64-frame table fit and prints that leg alone")
    parser.add_argument("--no-serving", action="store_true")
    args = parser.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(
            "bench.py measures one GPU: this build runs one system per GPU and has no multi-GPU "
            "path (spatial decomposition across GPUs is a design note, not code). Use --gpus 1.",
            file=sys.stderr,
        )
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    if args.kernel == "copy":
        source = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        target = torch.empty_like(source)
        source.fill_(1.0)
        for _ in range(20):
            target.copy_(source)
        torch.cuda.synchronize()
        print(json.dumps({"kernel": "copy", "bytes_read": source.numel() * 4,
                          "bytes_written": source.numel() * 4, "launches": 20}))
        return 0
    if args.kernel == "md-batch":
        print(json.dumps({"kernel": "md-batch", **bench_md_batch(torch, plain=False)}))
        return 0
    if args.kernel == "published":
        print(json.dumps({"kernel": "published", "sizes": bench_published_sizes(torch)}))
        return 0
    if args.kernel == "fitting":
        print(json.dumps({"kernel": "fitting", **bench_fitting(torch)}))
        return 0
    if args.kernel == "sum":
        from nabladyn_amd import workloads

        x, cell = workloads.argon_lattice(PER_AXIS, seed=0)
        print(json.dumps({"kernel": "sum", **bench_sum(torch, x, cell)}))
        return 0
    log("building the 2,097,152-atom system ...")
    # the force kernel's counters: the schedule the autotuner takes for this system (4 lanes)
    sim, x, cell = build_simulation(torch, group=4 if args.kernel == "force" else None)
    if args.kernel == "force":
        # the liquid, not the first frame's lattice: 100 steps first
        sim.run(100)
        torch.cuda.synchronize()
        seconds = time_force_kernel(torch, sim, launches=20)
        pairs = sim._neighbors.num_pairs
        print(json.dumps({"kernel": "force", "launches": 23, "steps_before": 100,
                          "ms": seconds * 1e3, "pairs_within_radius": pairs,
                          "group": sim.group}))
        return 0
    n = x.shape[0]
    # The step's cost includes its share of the list rebuilds the device decides (about one
    # every 29 steps here, 1.9 ms each) and of the pruned list's refreshes: a window shorter
    # than a few rebuild cycles would time mostly rebuild-free steps. So the window spans at
    # least MIN_TIMED_STEPS steps (about 30 rebuilds, each counted and reported), after at least
    # MIN_WARMUP_STEPS steps that melt the starting lattice into the liquid whose rebuild cadence
    # is measured.
    steps = max(args.steps, MIN_TIMED_STEPS)
    warmup = max(args.warmup, MIN_WARMUP_STEPS)
    if (steps, warmup) != (args.steps, args.warmup):
        log(f"timing {steps} steps after {warmup} of warm-up (asked for {args.steps} after "
            f"{args.warmup}): the window must hold the device-decided rebuilds")
    torch.cuda.synchronize()
    log(f"warm-up {warmup} steps ...")
    sim.run(warmup)
    torch.cuda.synchronize()
    rebuilds_before = sim._neighbors._rebuild_count
    refreshes_before = sim._neighbors.refresh_count if sim._neighbors.inner_skin else 0
    pairs_before = sim._neighbors.num_pairs
    log(f"timing {steps} steps ...")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim._replay(steps)  # 25-step graphs, then single steps
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sim._step += steps
    sim._check(sim._step)
    rebuilds = sim._neighbors._rebuild_count - rebuilds_before
    refreshes = (sim._neighbors.refresh_count - refreshes_before) if sim._neighbors.inner_skin else 0
    pairs_after = sim._neighbors.num_pairs
    # measured only: None when the window held no rebuild
    steps_per_rebuild = steps / rebuilds if rebuilds else None
    value = n * steps / elapsed
    ms_per_step = 1e3 * elapsed / steps
    pairs = 0.5 * (pairs_before + pairs_after)
    log(f"{ms_per_step:.4f} ms/step, {rebuilds} rebuilds and {refreshes} pruned-list refreshes "
        f"in {steps} steps, {pairs:.0f} pairs within R")

    # the dominant kernel: forces over the full list + second kick
    t_force = time_force_kernel(torch, sim)
    s = 4
    # Algorithmic bytes of one force launch, the reference algorithm's traffic: its half list
    # (two int32 per pair within R), the atom's own position and type, its count, F written, v
    # read and written, m and 1/m.
    force_bytes = 8 * pairs + n * (4 * s + 4 + 3 * s + 6 * s + 2 * s)
    achieved = force_bytes / t_force
    # the per-step algorithmic bytes of the whole step, the rebuild share at the measured cadence
    step_bytes = None
    if steps_per_rebuild is not None:
        step_bytes = (8 * pairs + n * (21 * s + 48)
                      + (8 * pairs + n * (9 * s + 20)) / steps_per_rebuild)
    traffic = None
    traffic_file = ROOT / "profiles" / "force_kernel_traffic.json"
    if traffic_file.exists():
        traffic = json.loads(traffic_file.read_text()).get("bytes_per_launch")
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "steps_requested": args.steps,
        "warmup_requested": args.warmup,
        "rebuilds_timed": rebuilds,
        "refreshes_timed": refreshes,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded jittered lattice, Maxwell-Boltzmann velocities at 120 K)",
        "config": {
            "workload": "liquid-argon LJ NVT, 2,097,152 atoms (128^3 jittered lattice), fp32",
            "atoms": n,
            "box_nm": float(cell[0]),
            "cutoff_nm": CUTOFF,
            "skin_nm": SKIN,
            "timestep_ps": 0.004,
            "thermostat": "Nose-Hoover chain of 3, 120 K, 0.1 ps",
            "pairs_within_radius": pairs,
            "steps_per_rebuild": steps_per_rebuild,
            "force_schedule_lanes_per_row": sim.group,
            "pruned_list_skin_nm": sim._neighbors.inner_skin,
            "steps_per_pruned_refresh": steps / refreshes if refreshes else None,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "force_kernel (one block per brick, halo in LDS, pruned rows; LJ forces + second half kick)",
            # algorithmic bytes (the reference algorithm's) per launch / event-timed duration
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "kernel_ms": t_force * 1e3,
            "algorithmic_bytes_per_launch": force_bytes,
            # the bytes the kernel really moves (rocprofv3 FETCH_SIZE + WRITE_SIZE per launch,
            # scaled by a copy kernel's reading) over the same duration
            "measured_GBps": traffic / t_force / 1e9 if traffic else None,
            "measured_frac": traffic / t_force / HBM_PEAK if traffic else None,
        },
        "step_roofline": {
            "algorithmic_bytes_per_step": step_bytes,
            "steps_per_rebuild": steps_per_rebuild,
            "achieved_GBps": step_bytes * steps / elapsed / 1e9 if step_bytes else None,
            "frac": step_bytes * steps / elapsed / HBM_PEAK if step_bytes else None,
        },
    }
    if not args.no_batching:
        log("batching: 256 ragged systems ...")
        result["batching"] = bench_batching(torch)
        log("batched MD: the serving stream's 14 NVT systems ...")
        result["md_batching"] = bench_md_batch(torch)
        log("potential.sum on the 2M frame ...")
        result["sum"] = bench_sum(torch, x, cell)
        log("the reference's published sizes (512 .. 32,768 atoms) ...")
        result["published_sizes"] = bench_published_sizes(torch)
        log("fitting: 64 frames of 4,096 beads, a 200-knot table ...")
        result["fitting"] = bench_fitting(torch)
    if not args.no_serving:
        log("serving: a Poisson request stream ...")
        result["serving"] = bench_serving(torch)
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle on the host cores) ...")
        del sim
        torch.cuda.empty_cache()
        result["cpu_baseline"] = cpu_baseline(x, cell, steps_per_rebuild)
        result["cpu_baseline"].update(host=cpu_host(), small_configs=cpu_small_configs())
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
