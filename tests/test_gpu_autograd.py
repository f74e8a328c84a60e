"""Differentiable models on the GPU: ``forward``'s energy under ``torch.func`` and
``torch.autograd``, at zero and nonzero strain, and the parameter gradients, against the oracle's
fp64 restatement of the reference's energies differentiated by autograd
(``potential/_base.py:82-106``, ``quantity.py:93-117``).

Tolerances, fp64 (relative to the largest entry of the oracle's result): E 1e-12, forces and
virials 1e-11, parameter gradients 1e-10 (sums of many fp64 terms added with atomics in another
order). fp32: forces from ``torch.func.grad(model)`` equal ``quantity``'s to 2e-6 of max|F|
(both are the same kernel's forces, added with atomics in another order).
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import potential, quantity
from nabladyn_amd.neighbors import NeighborList
from oracle import neighbors as on
from oracle import potentials as op

pytestmark = pytest.mark.gpu

N, L = 1500, 4.0
STRAIN = [[0.012, 0.003, -0.002], [0.001, -0.008, 0.004], [-0.003, 0.002, 0.006]]


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _system(device, dtype, kind, seed=3):
    rng = np.random.default_rng(seed)
    x = on.jittered_lattice(12, L, 0.06, seed=seed)[:N] + rng.uniform(-0.2, 0.2, (N, 3))
    types = torch.tensor(rng.integers(0, 2, N), device=device)
    if kind == "lj":
        model = potential.lennard_jones(
            atom_types=types, sigma=torch.tensor([0.30, 0.33, 0.35], dtype=dtype, device=device),
            epsilon=torch.tensor([0.7, 0.9, 1.1], dtype=dtype, device=device), cutoff=1.1,
            trainable=("sigma", "epsilon"))
    else:
        r = 0.25 + np.arange(120) * (1.1 - 0.25) / 119
        table = np.stack([4 * e * ((0.33 / r) ** 12 - (0.33 / r) ** 6) for e in (0.8, 1.0, 0.6)])
        table = np.clip(table, None, 60.0) + rng.normal(0, 0.05, table.shape)
        table[:, -1] = 0
        model = potential.tabulated(atom_types=types, table=torch.tensor(table, dtype=dtype, device=device),
                                    start=torch.full((3,), 0.25, dtype=dtype, device=device),
                                    cutoff=1.1, trainable=("table",))
    xd = torch.tensor(x, dtype=dtype, device=device)
    cell = torch.full((3,), L, dtype=dtype, device=device)
    nl = NeighborList(model.cutoff, skin=0.1)
    nl.allocate(xd, cell)
    return model, xd, cell, nl.tensors()


def _oracle(model, x, cell, pairs, strain):
    """fp64 E and its gradients in positions, strain and parameters (the oracle's torch
    restatement of the reference's energy, autograd)."""
    xw = torch.tensor(on.wrap(x.cpu().double().numpy(), cell.cpu().double().numpy())[0],
                      requires_grad=True)
    c = cell.cpu().double()
    s = torch.tensor(strain, dtype=torch.float64, requires_grad=True)
    idx = pairs.index.cpu().long()
    types = model.atom_types.cpu()
    if isinstance(model, potential.LennardJones):
        params = [model.sigma.detach().cpu().double().requires_grad_(True),
                  model.epsilon.detach().cpu().double().requires_grad_(True)]
        e = op.t_lj_energy(xw, s, c, idx, types, *params, model.cutoff)
    else:
        params = [model.table.detach().cpu().double().requires_grad_(True)]
        e = op.t_table_energy(xw, s, c, idx, types, params[0], model.start.cpu().double(),
                              model.spacing.cpu().double(), model.cutoff)
    grads = torch.autograd.grad(e, [xw, s, *params])
    return float(e.detach()), [g.numpy() for g in grads]


@pytest.mark.parametrize("kind", ["lj", "table"])
@pytest.mark.parametrize("strained", [False, True])
def test_energy_gradients_match_the_oracle_fp64(device, kind, strained):
    model, x, cell, pairs = _system(device, torch.float64, kind)
    strain = torch.tensor(STRAIN if strained else np.zeros((3, 3)), dtype=torch.float64,
                          device=device)
    params = [p for p in model._differentiable_parameters() if p is not None]
    xr = x.clone().requires_grad_(True)
    sr = strain.clone().requires_grad_(True)
    e = model(xr, cell, pairs, sr)
    grads = torch.autograd.grad(e, [xr, sr, *params])
    e0, want = _oracle(model, x, cell, pairs, strain.cpu().numpy())
    assert abs(float(e) - e0) <= 1e-12 * abs(e0)
    assert _rel(grads[0].cpu(), want[0]) <= 1e-11
    assert _rel(grads[1].cpu(), want[1]) <= 1e-11
    for got, ref in zip(grads[2:], want[2:]):
        assert _rel(got.cpu(), ref) <= 1e-10
    if kind == "table":
        assert float(grads[2][:, -1].abs().max()) == 0.0  # the last knot reads as 0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_func_grad_of_the_model_equals_quantity(device, dtype):
    model, x, cell, pairs = _system(device, dtype, "lj")
    e, f, w = quantity.energy_force_virial(model)(x, cell, pairs)
    zero = torch.zeros((3, 3), dtype=dtype, device=device)
    (gx, gs), value = torch.func.grad_and_value(
        lambda p, s: model(p, cell, pairs, s), argnums=(0, 1))(x, zero)
    tol = 2e-6 if dtype == torch.float32 else 1e-13
    assert abs(float(value) - float(e)) <= tol * abs(float(e))
    assert float((-gx - f).abs().max()) <= tol * float(f.abs().max())
    assert float((-gs - w).abs().max()) <= tol * float(w.abs().max())


def test_quantity_of_a_plain_callable_matches_the_oracle(device):
    """A user energy (the LJ model plus a harmonic restraint pulling atom 0 to a point) goes
    through torch.func as in the reference."""
    model, x, cell, pairs = _system(device, torch.float64, "lj")
    anchor = x[0].clone() + 0.05

    def energy(positions, cell_, pairs_, strain=None):
        return model(positions, cell_, pairs_, strain) + 7.5 * (positions[0] - anchor).square().sum()

    e, f, w = quantity.energy_force_virial(energy)(x, cell, pairs)
    e0, want = _oracle(model, x, cell, pairs, np.zeros((3, 3)))
    restraint = 7.5 * float((x[0] - anchor).square().sum())
    assert abs(float(e) - (e0 + restraint)) <= 1e-12 * abs(e0)
    f_ref = -want[0]
    f_ref[0] += (-15.0 * (x[0] - anchor)).cpu().numpy()
    assert _rel(f.cpu(), f_ref) <= 1e-11
    assert _rel(w.cpu(), -want[1]) <= 1e-11


def test_vmap_over_frames_and_second_derivative_refused(device):
    model, x, cell, pairs = _system(device, torch.float64, "lj")
    frames = torch.stack([x, x + 1e-3, x - 2e-3])
    energies = torch.func.vmap(lambda p: model(p, cell, pairs))(frames)
    one_by_one = torch.stack([model(p, cell, pairs) for p in frames])
    assert torch.equal(energies, one_by_one)
    xr = x.clone().requires_grad_(True)
    (g,) = torch.autograd.grad(model(xr, cell, pairs), xr, create_graph=True)
    with pytest.raises(RuntimeError, match="second derivatives"):
        torch.autograd.grad(g.square().sum(), model.sigma)
