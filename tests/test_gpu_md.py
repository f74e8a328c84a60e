"""The MD step (one HIP graph) against the oracle's eager integrator and the reference's
published NVT conserved-energy band.

Tolerances: fp64 trajectories match the oracle's to 1e-9 (relative, max norm) over 20 steps;
fp32 to 1e-4 over 10 steps (chaotic growth of fp32 rounding differences, which starts at the
1e-7 level of the forces). NVE drift and the NVT conserved energy are compared with the
reference's documented figures.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import md, potential
from nabladyn_amd.md._simulation import INNER_SKIN
from oracle import md as omd
from oracle import neighbors as on
from _parity import check_against_route, table_route

pytestmark = pytest.mark.gpu

MASS = 39.948


def _lj(n, dtype, device):
    return potential.lennard_jones(
        atom_types=torch.zeros(n, dtype=torch.int64, device=device),
        sigma=torch.tensor([0.34], dtype=dtype, device=device),
        epsilon=torch.tensor([0.996], dtype=dtype, device=device), cutoff=1.0)


def _config1(dtype, device, seed_pos=0, seed_vel=1):
    x = on.jittered_lattice(8, 2.90, 0.02, seed=seed_pos)
    masses = torch.full((512,), MASS, dtype=dtype)
    v = md.maxwell_boltzmann(masses, 120.0, seed_vel)
    return (torch.tensor(x, dtype=dtype), v, masses, torch.full((3,), 2.90, dtype=dtype))


def test_maxwell_boltzmann_matches_restatement():
    masses = torch.full((512,), MASS, dtype=torch.float64)
    a = md.maxwell_boltzmann(masses, 120.0, 1)
    b = omd.maxwell_boltzmann(masses, 120.0, 1)
    assert torch.equal(a, b)


def test_first_frame_matches_quantity(device):
    """The MD path's full-list evaluation equals the half-list kernel on the same frame."""
    from nabladyn_amd import quantity
    from nabladyn_amd.neighbors import NeighborList

    dtype = torch.float64
    x, v, m, L = _config1(dtype, device)
    model = _lj(512, dtype, device)
    sim = md.nve(model, x.to(device), v.to(device), m.to(device), L.to(device), timestep=0.004)
    nl = NeighborList(1.0, skin=0.1)
    nl.allocate(x.to(device), L.to(device))
    e, f, w = quantity.energy_force_virial(model)(x.to(device), L.to(device), nl.tensors())
    ferr = float((sim._to_atoms(sim.forces) - f).abs().max() / f.abs().max())
    assert ferr < 1e-12, ferr
    assert abs(float(sim.energy) - float(e)) < 1e-10 * abs(float(e))
    assert float((sim.virial - w).abs().max()) < 1e-10 * float(w.abs().max())


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_single_steps_match_oracle(device, dtype):
    """Step by step (NVE), positions, velocities and forces against the eager restatement."""
    x, v, m, L = _config1(dtype, device)
    sim = md.nve(_lj(512, dtype, device), x.to(device), v.to(device), m.to(device), L.to(device),
                 timestep=0.004)
    ref = omd.Integrator(omd.LJModel(np.zeros(512, np.int64), [0.34], [0.996], 1.0, dtype),
                         x, v, m, L, dt=0.004, skin=0.1)
    tol = 1e-11 if dtype == torch.float64 else 2e-5
    report = []
    for step in range(1, 6):
        ref.step()
        sim._launch_step()
        torch.cuda.synchronize()
        state = sim.state()
        ex = float((state.positions.cpu() - ref.x).abs().max())
        ev = float((state.velocities.cpu() - ref.v).abs().max() / ref.v.abs().max())
        ef = float((state.forces.cpu() - ref.forces).abs().max() / ref.forces.abs().max())
        report.append((step, ex, ev, ef, sim._neighbors._rebuild_count, ref.list.rebuilds))
    assert all(r[1] < tol and r[2] < 10 * tol and r[3] < 10 * tol for r in report), report


@pytest.mark.parametrize("dtype,steps,tol", [(torch.float64, 20, 1e-9), (torch.float32, 10, 1e-4)])
@pytest.mark.parametrize("ensemble", ["nve", "nvt"])
def test_trajectory_matches_oracle(device, dtype, steps, tol, ensemble):
    x, v, m, L = _config1(dtype, device)
    model = _lj(512, dtype, device)
    kw = dict(timestep=0.004, skin=0.1)
    if ensemble == "nvt":
        kw.update(temperature=120.0, thermostat_timescale=0.1)
        sim = md.nvt(model, x.to(device), v.to(device), m.to(device), L.to(device), **kw)
    else:
        sim = md.nve(model, x.to(device), v.to(device), m.to(device), L.to(device), **kw)
    ref = omd.Integrator(omd.LJModel(np.zeros(512, np.int64), [0.34], [0.996], 1.0, dtype),
                         x, v, m, L, dt=0.004, skin=0.1,
                         temperature=120.0 if ensemble == "nvt" else None,
                         tau=0.1 if ensemble == "nvt" else None)
    for _ in range(steps):
        ref.step()
    sim.run(steps)
    state = sim.state()
    pos = state.positions.cpu().double() + state.images.cpu().double() * L.double()
    ref_pos = ref.x.double() + ref.images.double() * L.double()
    scale = float(ref_pos.abs().max())
    assert float((pos - ref_pos).abs().max()) / scale <= tol
    vel_err = float((state.velocities.cpu().double() - ref.v.double()).abs().max())
    assert vel_err / float(ref.v.abs().max()) <= 10 * tol
    f_err = float((state.forces.cpu().double() - ref.forces.double()).abs().max())
    assert f_err / float(ref.forces.abs().max()) <= 10 * tol
    if ensemble == "nvt":
        # K is summed in fp64 here and in the frame's dtype by the reference's eager torch, so in
        # fp32 the chain's driving term differs at the 1e-7 level
        rtol, atol = (1e-8, 1e-12) if dtype == torch.float64 else (1e-3, 1e-6)
        assert torch.allclose(state.thermostat.positions.cpu(), ref.chain_x, rtol=rtol, atol=atol)
        assert torch.allclose(state.thermostat.momenta.cpu(), ref.chain_p, rtol=rtol, atol=atol)


def test_nvt_conserved_energy_band_argon(argon, device):
    """The reference's published NVT run (examples/README.md:58-64): argon_512, fp32, 2,500
    steps of 4 fs at 120 K, tau 0.1 ps: "the conserved energy stays within 0.03 kJ/mol of
    -1,713.4 kJ/mol", mean T 119.1 K. The centre is printed to 0.1 kJ/mol, so the statement
    allows values in [-1713.48, -1713.32] with a spread of at most 0.06; fp32 runs that round
    differently (each schedule here) wander differently inside that band."""
    dtype = torch.float32
    x = torch.tensor(argon["positions"], dtype=dtype, device=device)
    v = torch.tensor(argon["velocities"], dtype=dtype, device=device)
    L = torch.tensor(argon["cell"], dtype=dtype, device=device)
    m = torch.full((512,), MASS, dtype=dtype, device=device)
    sim = md.nvt(_lj(512, dtype, device), x, v, m, L, timestep=0.004, temperature=120.0,
                 thermostat_timescale=0.1)
    samples = sim.run(2500, sample_every=100)
    conserved = samples["conserved_energy"].numpy()
    first = sim  # keep
    assert abs(conserved[0] - (-1713.4)) < 0.05
    assert np.abs(conserved - (-1713.4)).max() <= 0.08
    assert np.ptp(conserved) <= 0.06
    assert abs(samples["temperature"].mean() - 119.1) < 3.0
    assert sim._neighbors._rebuild_count > 10


@pytest.mark.parametrize("dtype,bound", [(torch.float64, 2e-4), (torch.float32, 5e-3)])
def test_nve_energy_drift(device, dtype, bound):
    x, v, m, L = _config1(dtype, device)
    sim = md.nve(_lj(512, dtype, device), x.to(device), v.to(device), m.to(device), L.to(device),
                 timestep=0.004)
    samples = sim.run(1000, sample_every=50)
    total = samples["total_energy"].numpy()
    # per-atom energy excursion over 4 ps
    assert np.abs(total - total[0]).max() / 512 <= bound


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("n,length", [(512, 2.9), (13824, 8.7), (4096, 5.8)])
def test_md_full_list_pair_set(device, dtype, n, length):
    """The MD list (sub-bins, culled stencil, slot order) holds exactly the pairs within R of
    the oracle, in both orientations, at its build and after rebuilds during a run."""
    npdt = np.float32 if dtype == torch.float32 else np.float64
    per = round(n ** (1 / 3))
    x = on.jittered_lattice(per, length, 0.02, seed=5)
    masses = torch.full((n,), MASS, dtype=dtype)
    v = md.maxwell_boltzmann(masses, 300.0, 2)
    L = torch.full((3,), length, dtype=dtype)
    sim = md.nve(_lj(n, dtype, device), torch.tensor(x, dtype=dtype, device=device),
                 v.to(device), masses.to(device), L.to(device), timestep=0.004, skin=0.2)

    def check():
        pairs = sim._neighbors.atom_pairs().numpy()
        xw = sim.state().positions.cpu().numpy()
        # the list was built at the last rebuild's positions: compare at those
        ref_pos = sim._to_atoms(sim._neighbors.reference_positions).cpu().numpy()
        brute = on.brute_pairs(ref_pos, L.numpy().astype(npdt), 1.2)
        sym = np.concatenate([brute, brute[:, ::-1]])
        key = lambda a: a[np.lexsort((a[:, 1], a[:, 0]))]
        np.testing.assert_array_equal(key(pairs), key(sym))
        return xw

    check()
    before = sim._neighbors._rebuild_count
    sim.run(60)
    assert sim._neighbors._rebuild_count > before
    check()


@pytest.mark.parametrize("inner_skin", [0.05, INNER_SKIN])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_md_pruned_list_pair_set(device, dtype, inner_skin):
    """The pruned list the force kernel reads holds exactly the full list's pairs within
    cutoff + inner skin at its last refresh (both orientations), is refreshed before any atom
    moves half that skin, and so always holds every pair within the cutoff. At the default
    pruned skin (0.10 nm inside a 0.2 nm list skin) as the 2M-atom step runs it, and at 0.05."""
    n, length = 4096, 5.8
    npdt = np.float32 if dtype == torch.float32 else np.float64
    x = on.jittered_lattice(16, length, 0.02, seed=6)
    masses = torch.full((n,), MASS, dtype=dtype)
    v = md.maxwell_boltzmann(masses, 300.0, 3)
    L = torch.full((3,), length, dtype=dtype)
    sim = md.Simulation(_lj(n, dtype, device), torch.tensor(x, dtype=dtype, device=device),
                        v.to(device), masses.to(device), L.to(device), timestep=0.004, skin=0.2,
                        capacity_multiplier=1.25, images=None, step=0, inner_skin=inner_skin)
    key = lambda a: a[np.lexsort((a[:, 1], a[:, 0]))]
    nb = sim._neighbors
    refreshes, rebuilds = [], []
    for _ in range(8):
        pruned = nb.atom_pairs(pruned=True).numpy()
        full = {tuple(p) for p in nb.atom_pairs().numpy()}
        at_refresh = sim._to_atoms(nb.buffers["prune_positions"]).cpu().numpy()
        brute = on.brute_pairs(at_refresh, L.numpy().astype(npdt), 1.0 + inner_skin)
        expected = np.array([p for p in np.concatenate([brute, brute[:, ::-1]])
                             if (int(p[0]), int(p[1])) in full])
        np.testing.assert_array_equal(key(pruned), key(expected))
        now = on.brute_pairs(sim.state().positions.cpu().numpy(), L.numpy().astype(npdt), 1.0)
        have = {tuple(p) for p in pruned}
        assert all((int(i), int(j)) in have for i, j in now)
        refreshes.append(nb.refresh_count)
        rebuilds.append(nb._rebuild_count)
        sim.run(7)
    # every build refreshes the pruned list; some refreshes come between builds
    assert refreshes[-1] - refreshes[0] > rebuilds[-1] - rebuilds[0], (refreshes, rebuilds)


def test_graph_replay_equals_eager_launches(device):
    x, v, m, L = _config1(torch.float32, device)
    # one fixed force schedule for both: the autotuner's timing-based choice may differ between
    # two simulations, and a different lane group sums the forces in a different order
    kw = dict(timestep=0.004, skin=0.1, capacity_multiplier=1.25, images=None, step=0,
              temperature=120.0, thermostat_timescale=0.1, force_group=8)
    a = md.Simulation(_lj(512, torch.float32, device), x.to(device), v.to(device), m.to(device),
                      L.to(device), **kw)
    b = md.Simulation(_lj(512, torch.float32, device), x.to(device), v.to(device), m.to(device),
                      L.to(device), **kw)
    a.run(50)
    for _ in range(50):
        b._launch_step()
    b._step += 50
    sa, sb = a.state(), b.state()
    assert torch.equal(sa.positions, sb.positions)
    assert torch.equal(sa.velocities, sb.velocities)
    assert torch.equal(sa.images, sb.images)


def test_state_continues_run(device):
    x, v, m, L = _config1(torch.float64, device)
    model = _lj(512, torch.float64, device)
    kw = dict(timestep=0.004, temperature=120.0, thermostat_timescale=0.1)
    a = md.nvt(model, x.to(device), v.to(device), m.to(device), L.to(device), **kw)
    a.run(30)
    s = a.state()
    b = md.nvt(model, s.positions, s.velocities, m.to(device), s.cell, images=s.images,
               step=s.step, thermostat=s.thermostat, **kw)
    a.run(10)
    b.run(10)
    sa, sb = a.state(), b.state()
    # b rebuilt its list at the restart, so its force sums may round differently
    assert float((sa.positions - sb.positions).abs().max()) < 1e-9
    assert sa.step == sb.step == 40


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("num_types", [1, 2])
def test_md_table_forces_vs_oracle(device, dtype, num_types):
    """The MD force kernel with a 200-knot table (one type: its segments staged in LDS; two:
    gathered per pair) on the same pairs as the reference: fp32 max|dF| / max|F|,
    max|dW| / max|W| and |dE| / max(1, |E|) within 1e-5 of the reference's own fp32 route, the
    fp64 errors printed beside; fp64 within 3e-12 of the oracle's closed form. Random
    positions, so some pairs fall below the first knot (the tangent branch)."""
    from oracle import potentials as op

    rng = np.random.default_rng(5)
    n = 1500
    cell = np.array([4.2, 3.9, 4.5])
    x = rng.uniform(0, 1, (n, 3)) * cell
    rows = num_types * (num_types + 1) // 2
    r = 0.3 + np.arange(200) * (1.2 - 0.3) / 199
    lj = 4 * 0.996 * ((0.34 / r) ** 12 - (0.34 / r) ** 6)
    table = np.stack([np.clip(lj, None, 50.0) + rng.normal(0, 0.05, 200) for _ in range(rows)])
    table[:, -1] = 0.0
    types = rng.integers(0, num_types, n)
    model = potential.tabulated(
        atom_types=torch.tensor(types, dtype=torch.int64, device=device),
        table=torch.tensor(table, dtype=dtype, device=device),
        start=torch.full((rows,), 0.3, dtype=dtype, device=device), cutoff=1.2)
    xd = torch.tensor(x, dtype=dtype, device=device)
    zeros = torch.zeros_like(xd)
    masses = torch.full((n,), MASS, dtype=dtype, device=device)
    sim = md.nve(model, xd, zeros, masses, torch.tensor(cell, dtype=dtype, device=device),
                 timestep=0.002, skin=0.2)
    xw = on.wrap(xd.cpu().numpy(), cell)[0]
    pairs = on.brute_pairs(xw, cell, 1.2).T.copy()
    E0, F0, W0 = op.table_efw(xw, cell, pairs, types, table.astype(np.float64) if dtype == torch.float64
                              else table.astype(np.float32).astype(np.float64),
                              np.full(rows, 0.3), 1.2, spacing=model.spacing.cpu().numpy())
    F = sim._to_atoms(sim.forces).cpu().numpy()
    W = sim.virial.cpu().numpy()
    if dtype == torch.float64:
        tol = 3e-12
        assert np.abs(F - F0).max() <= tol * np.abs(F0).max()
        assert np.abs(W - W0).max() <= tol * np.abs(W0).max()
        assert abs(float(sim.energy) - E0) <= tol * max(1.0, abs(E0))
        return
    route = table_route(xw, cell.astype(np.float32), pairs, types, table.astype(np.float32),
                        np.full(rows, 0.3, np.float32), model.spacing.cpu().numpy(), 1.2)
    check_against_route(f"MD table {num_types} types, fp32", (float(sim.energy), F, W), route,
                        (E0, F0, W0))
