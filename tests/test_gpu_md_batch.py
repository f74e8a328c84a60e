"""Batched MD (``md.SimulationBatch``): every member of a batch ends bitwise where the same
simulation run alone (``Simulation.run``, one captured graph per step) ends.

Members: fp32 and fp64 LJ argon and 200-knot table systems (by kind and mixed in one batch)
of 125 to 4,096 atoms in boxes of 2.5 to 5.9 nm at 8 to 21 nm^-3 (some boxes short against
2 x (cutoff + skin), one member without a pruned list, NVE and NVT mixed), stepped far enough
that each rebuilds its list several times. The bar is bitwise: a member's kernels get the blocks its own step launches and
run the same code on its data.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import md, potential, workloads
from oracle import neighbors as on

pytestmark = pytest.mark.gpu

MASS = 39.948


def _model(kind, n, dtype, device):
    types = torch.zeros(n, dtype=torch.int64, device=device)
    if kind == "lj":
        return potential.lennard_jones(
            atom_types=types, sigma=torch.tensor([0.34], dtype=dtype, device=device),
            epsilon=torch.tensor([0.996], dtype=dtype, device=device), cutoff=1.0)
    return potential.tabulated(
        atom_types=types, table=torch.tensor(workloads.lj_table()[None], dtype=dtype, device=device),
        start=torch.tensor([0.3], dtype=dtype, device=device), cutoff=1.2)


def _system(kind, per, length, dtype, device, seed, nvt=True, inner=True):
    x = on.jittered_lattice(per, length, 0.02, seed=seed)
    n = x.shape[0]
    masses = torch.full((n,), MASS, dtype=dtype, device=device)
    v = md.maxwell_boltzmann(masses, 120.0, seed + 100).to(dtype)
    kw = dict(timestep=0.004, skin=0.2, capacity_multiplier=1.25, images=None, step=0,
              force_group=8, inner_skin=md._simulation.INNER_SKIN if inner else None)
    if nvt:
        kw.update(temperature=120.0, thermostat_timescale=0.1)
    args = (_model(kind, n, dtype, device), torch.tensor(x, dtype=dtype, device=device), v,
            masses, torch.full((3,), length, dtype=dtype, device=device))
    return lambda graphs=True: md.Simulation(*args, graphs=graphs, **kw)


def _members(kind, dtype, device):
    # atoms per side and box edge (nm)
    shapes = [(5, 2.5), (8, 2.9), (11, 4.0), (16, 5.9), (6, 3.0)]
    return [_system(kind, per, length, dtype, device, seed=7 * k + 1, nvt=k % 2 == 0,
                    inner=k != 3)
            for k, (per, length) in enumerate(shapes)]


def _same(a, b):
    sa, sb = a.state(), b.state()
    return (torch.equal(sa.positions, sb.positions) and torch.equal(sa.velocities, sb.velocities)
            and torch.equal(sa.forces, sb.forces) and torch.equal(sa.images, sb.images)
            and sa.step == sb.step)


@pytest.mark.parametrize("kind,dtype", [("lj", torch.float32), ("table", torch.float32),
                                        ("lj", torch.float64)])
def test_batch_members_equal_each_alone(device, kind, dtype):
    makers = _members(kind, dtype, device)
    alone = [make() for make in makers]
    for s in alone:
        s.run(120)
    batched = [make(graphs=False) for make in makers]
    batch = md.SimulationBatch(batched)
    batch.advance(70)
    batch.advance(50)
    batch.check()
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert _same(a, b), (kind, str(dtype), k, a.state().positions.shape[0])
    # rebuilds happened inside the batch
    assert all(s._neighbors._rebuild_count > 1 for s in batched)


def test_batch_then_alone_continues(device):
    makers = _members("lj", torch.float32, device)[:3]
    reference = [make() for make in makers]
    for s in reference:
        s.run(80)
    sims = [make() for make in makers]
    md.SimulationBatch(sims).advance(40)
    for s in sims:
        s.run(40)
    for a, b in zip(reference, sims):
        assert _same(a, b)
        np.testing.assert_equal(a.sample()["temperature"], b.sample()["temperature"])


def test_batch_single_steps_equal_a_run(device):
    """A call of one step runs the prologue; later steps of a call are opened by the step
    before them. Either way each member ends bitwise where it ends alone."""
    makers = _members("table", torch.float32, device)[:3]
    alone = [make() for make in makers]
    for s in alone:
        s.run(45)
    batched = [make(graphs=False) for make in makers]
    batch = md.SimulationBatch(batched)
    for _ in range(15):
        batch.advance(1)
    batch.advance(30)
    batch.check()
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert _same(a, b), k


def test_mixed_potentials_in_one_batch(device):
    # LJ and one-type-table members share every launch but their force launches
    makers = _members("lj", torch.float32, device)[:3] + _members("table", torch.float32, device)[2:]
    alone = [make() for make in makers]
    for s in alone:
        s.run(90)
    batched = [make(graphs=False) for make in makers]
    batch = md.SimulationBatch(batched)
    force = md._batch.FORCE_LAUNCHES
    assert batch.grids[force["lj"]] > 0 and batch.grids[force["table"]] > 0
    batch.advance(90)
    batch.check()
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert _same(a, b), k


def _typed_system(kind, per, length, device, seed):
    x = on.jittered_lattice(per, length, 0.02, seed=seed)
    n = x.shape[0]
    rng = np.random.default_rng(seed)
    types = torch.tensor(rng.integers(0, 2, n), dtype=torch.int64, device=device)
    dtype = torch.float32
    if kind == "lj":
        model = potential.lennard_jones(
            atom_types=types, sigma=torch.tensor([0.34, 0.31, 0.29], dtype=dtype, device=device),
            epsilon=torch.tensor([0.996, 0.8, 0.6], dtype=dtype, device=device), cutoff=1.0)
    else:
        table = np.stack([workloads.lj_table(seed=seed + k) for k in range(3)])
        model = potential.tabulated(
            atom_types=types, table=torch.tensor(table, dtype=dtype, device=device),
            start=torch.full((3,), 0.3, dtype=dtype, device=device), cutoff=1.2)
    masses = torch.full((n,), MASS, dtype=dtype, device=device)
    v = md.maxwell_boltzmann(masses, 120.0, seed + 100).to(dtype)
    args = (model, torch.tensor(x, dtype=dtype, device=device), v, masses,
            torch.full((3,), length, dtype=dtype, device=device))
    kw = dict(timestep=0.004, skin=0.2, capacity_multiplier=1.25, images=None, step=0,
              force_group=4, temperature=120.0, thermostat_timescale=0.1)
    return lambda graphs=True: md.Simulation(*args, graphs=graphs, **kw)


def test_typed_potentials_in_one_batch(device):
    # two atom types: the typed LJ and typed table force launches in one batch
    makers = [_typed_system("lj", 8, 2.9, device, 3), _typed_system("table", 9, 3.3, device, 4),
              _typed_system("lj", 11, 4.0, device, 5), _typed_system("table", 8, 2.9, device, 6)]
    alone = [make() for make in makers]
    for s in alone:
        s.run(80)
    batched = [make(graphs=False) for make in makers]
    batch = md.SimulationBatch(batched)
    force = md._batch.FORCE_LAUNCHES
    assert batch.grids[force["typed_lj"]] > 0 and batch.grids[force["typed_table"]] > 0
    batch.advance(80)
    batch.check()
    for k, (a, b) in enumerate(zip(alone, batched)):
        assert _same(a, b), k


def test_batch_refuses_mixed_and_empty(device):
    lj = _members("lj", torch.float32, device)[0](graphs=False)
    wide = _members("lj", torch.float64, device)[0](graphs=False)
    with pytest.raises(ValueError):
        md.SimulationBatch([])
    with pytest.raises(ValueError):
        md.SimulationBatch([lj, wide])
    with pytest.raises(ValueError):
        md.SimulationBatch([lj, lj])
    table = _members("table", torch.float32, device)[0](graphs=False)
    assert md.batch_key(lj) == md.batch_key(table) != md.batch_key(wide)
    batch = md.SimulationBatch([lj])
    assert len(batch) == 1 and batch.grids[-1] == 1
    assert batch.grids[md._batch.FORCE_LAUNCHES["table"]] == 0
