"""The batched engine: every system's E, F, W against the oracle and the plain per-system path,
for every schedule; each system's results identical alone and in the mix.

Tolerance. fp32: max|dF| / max|F_ref|, |dW| / max|W_ref| and |dE| / max(1, |E_ref|) within 1e-5
of the reference's own fp32 route (its autograd evaluation, restated by the oracle, run in fp32
on the same positions and pairs), with each system's error against an fp64 evaluation printed
beside it. fp64: 1e-12 against the oracle's closed forms. The pairs come from the oracle's C
cell list (pinned by brute force in the oracle's own tests).

The full batching workload (256 systems, 64 .. 32,768 atoms, half LJ, half a 200-knot table)
runs once, fp32, autotuned schedule: every system against the route.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import batch, potential, workloads
from oracle import neighbors as on
from oracle import potentials as op
from _parity import ROUTE_TOL, lj_route, rel, table_route

pytestmark = pytest.mark.gpu


def _systems(device, dtype, count=24, high=4096, seed=2):
    raw = workloads.ragged_systems(count, high=high, seed=seed, min_length=2.45)
    return batch.make_systems(raw, device, dtype)


def _pairs(xw, cell, cutoff):
    """Unordered pairs within the cutoff, [2, P], from the oracle's cell list."""
    b = on.build(xw, cell, cutoff)
    return b.pairs[:, :b.pair_total]


def _reference(x, cell, model):
    """(fp64 closed form, the reference's route in the system's dtype) of E, F, W."""
    xw = on.wrap(x.cpu().numpy(), cell.cpu().numpy())[0]
    cn = cell.cpu().numpy()
    pairs = _pairs(xw, cn, model.cutoff)
    types = np.zeros(len(xw), np.int64)
    x64, c64 = xw.astype(np.float64), cn.astype(np.float64)
    if isinstance(model, potential.LennardJones):
        sigma, eps = model.sigma.detach().cpu().numpy(), model.epsilon.detach().cpu().numpy()
        fp64 = op.lj_efw(x64, c64, pairs, types, sigma, eps, model.cutoff)
        route = lj_route(xw, cn, pairs, types, sigma, eps, model.cutoff)
    else:
        args = (model.table.detach().cpu().numpy(), model.start.cpu().numpy())
        spacing = model.spacing.cpu().numpy()
        fp64 = op.table_efw(x64, c64, pairs, types, *args, model.cutoff, spacing=spacing)
        route = table_route(xw, cn, pairs, types, *args, spacing, model.cutoff)
    return fp64, route


def _check(k, got, fp64, route, dtype, label):
    e, f, w = got
    if dtype == torch.float64:
        assert rel(f, fp64[1]) <= 1e-12, (label, k)
        assert rel(w, fp64[2]) <= 1e-12, (label, k)
        assert abs(e - fp64[0]) <= 1e-12 * max(1.0, abs(fp64[0])), (label, k)
        return None
    errors = (rel(f, route[1]), rel(w, route[2]), abs(e - route[0]) / max(1.0, abs(route[0])),
              rel(f, fp64[1]), rel(w, fp64[2]))
    assert max(errors[:3]) <= ROUTE_TOL, (label, k, errors)
    return errors


def _evaluate_and_check(systems, b, dtype, label):
    energy, forces, virial = b.evaluate()
    label = label.format(chosen=getattr(b, "chosen", None))
    worst = np.zeros(5)
    for k, (f, (x, cell, model)) in enumerate(zip(b.split(forces), systems)):
        fp64, route = _reference(x, cell, model)
        got = (float(energy[k]), f.cpu().numpy(), virial[k].cpu().numpy())
        errors = _check(k, got, fp64, route, dtype, label)
        if errors is not None:
            worst = np.maximum(worst, errors)
    if dtype == torch.float32:
        print(f"{label}: worst over {len(systems)} systems: F vs route {worst[0]:.2e}, W vs route "
              f"{worst[1]:.2e}, E vs route {worst[2]:.2e}; F vs fp64 {worst[3]:.2e}, W vs fp64 "
              f"{worst[4]:.2e}")


@pytest.mark.parametrize("schedule", batch.SCHEDULES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_matches_oracle(device, schedule, dtype):
    systems = _systems(device, dtype, count=12 if dtype == torch.float64 else 24)
    b = batch.Batch(systems, schedule=schedule)
    _evaluate_and_check(systems, b, dtype, f"schedule {schedule}")


def test_batch_full_workload_256_systems(device):
    """The batching workload at its stated size: 256 systems of 64 .. 32,768 atoms (seed 2),
    fp32, the schedule autotuned; every system against the reference's route."""
    raw = workloads.ragged_systems(256, seed=2, min_length=2.45)
    systems = batch.make_systems(raw, device)
    assert max(int(x.shape[0]) for x, _, _ in systems) > 16384
    b = batch.Batch(systems, schedule="auto")
    _evaluate_and_check(systems, b, torch.float32, "256 systems, schedule {chosen}")


def test_batch_matches_plain_path(device):
    systems = _systems(device, torch.float32)
    b = batch.Batch(systems, schedule=("wave", 2))
    energy, forces, virial = b.evaluate()
    plain = batch.plain_efw(systems)
    for k, (f, (e1, f1, w1)) in enumerate(zip(b.split(forces), plain)):
        assert rel(f.cpu(), f1.cpu()) <= 2e-5, k
        assert rel(virial[k].cpu(), w1.cpu()) <= 2e-5, k


@pytest.mark.parametrize("schedule", [("wave", 2), ("thread", 1)])
def test_alone_equals_in_the_mix(device, schedule):
    systems = _systems(device, torch.float32)
    mix = batch.Batch(systems, schedule=schedule)
    e_mix, f_mix, w_mix = (t.clone() for t in mix.evaluate())
    f_mix = mix.split(f_mix)
    for k in (0, 5, 11, 23):
        alone = batch.Batch([systems[k]], schedule=schedule)
        e, f, w = alone.evaluate()
        assert torch.equal(e[0], e_mix[k])
        assert torch.equal(f, f_mix[k])
        assert torch.equal(w[0], w_mix[k])


def test_autotuner_picks_a_schedule(device):
    systems = _systems(device, torch.float32, count=16)
    b = batch.Batch(systems)
    b.evaluate()
    assert b.chosen in batch.SCHEDULES
    assert set(b.timings) == set(batch.SCHEDULES)
