"""Host logic of the differentiable models, on the CPU: the autograd function around the fused
pair kernel (``potential/_energy.py``) composes with ``torch.func.grad``, ``torch.autograd``,
``vmap`` and ``quantity``'s generic route, and refuses a second derivative.

The kernel launches are replaced here by the oracle's torch restatement of the same energy (its
E, F, W at the strain, and its parameter gradient by autograd), so only the plumbing is under
test; the GPU tests check the kernels against the oracle.
"""

import numpy as np
import pytest
import torch

from nabladyn_amd import potential, quantity
from nabladyn_amd.neighbors import NeighborPairs
from nabladyn_amd.potential import _energy
from oracle import neighbors as on
from oracle import potentials as op

N, L = 200, 2.6


def _oracle_energy(model, positions, strain, cell, index):
    idx = index.long()
    if isinstance(model, potential.LennardJones):
        return op.t_lj_energy(positions, strain, cell, idx, model.atom_types, model.sigma,
                              model.epsilon, model.cutoff)
    return op.t_table_energy(positions, strain, cell, idx, model.atom_types, model.table,
                             model.start, model.spacing, model.cutoff)


def _fake_evaluate(model, positions, cell, index, strain, energy=True, forces=True, virial=True):
    x = positions.detach()
    s = strain.detach() if strain is not None else x.new_zeros((3, 3))
    (gx, gs), e = torch.func.grad_and_value(
        lambda p, q: _oracle_energy(model, p, q, cell.detach(), index), argnums=(0, 1))(x, s)
    return e.detach(), -gx.detach(), -gs.detach()


def _fake_parameter_gradient(model, positions, cell, index, strain):
    s = strain.detach() if strain is not None else positions.new_zeros((3, 3))
    params = [p for p in model._differentiable_parameters() if p is not None]
    copies = [p.detach().clone().requires_grad_(True) for p in params]
    with torch.enable_grad():  # called from a backward pass
        if isinstance(model, potential.LennardJones):
            e = op.t_lj_energy(positions.detach(), s, cell, index.long(), model.atom_types,
                               copies[0], copies[1], model.cutoff)
        else:
            e = op.t_table_energy(positions.detach(), s, cell, index.long(), model.atom_types,
                                  copies[0], model.start, model.spacing, model.cutoff)
        grads = torch.autograd.grad(e, copies)
    if isinstance(model, potential.LennardJones):
        return torch.stack(grads).double()
    return grads[0].double()


@pytest.fixture
def fake_kernels(monkeypatch):
    monkeypatch.setattr(_energy, "evaluate", _fake_evaluate)
    monkeypatch.setattr(_energy, "parameter_gradient", _fake_parameter_gradient)


def _frame(dtype=torch.float64, seed=0):
    x = on.jittered_lattice(6, L, 0.05, seed=seed)[:N]
    pairs = on.brute_pairs(x, np.full(3, L), 1.0).T.astype(np.int32)
    return (torch.tensor(x, dtype=dtype), torch.full((3,), L, dtype=dtype),
            NeighborPairs(torch.tensor(pairs), N, 1.0))


def _models(dtype=torch.float64):
    types = torch.tensor(np.arange(N) % 2, dtype=torch.int64)
    lj = potential.lennard_jones(atom_types=types, sigma=torch.tensor([0.30, 0.32, 0.34], dtype=dtype),
                                 epsilon=torch.tensor([0.8, 0.9, 1.0], dtype=dtype), cutoff=1.0,
                                 trainable=("sigma", "epsilon"))
    r = 0.3 + np.arange(40) * 0.7 / 39
    table = np.stack([4 * ((0.33 / r) ** 12 - (0.33 / r) ** 6) for _ in range(3)])
    table[:, -1] = 0
    tab = potential.tabulated(atom_types=types, table=torch.tensor(table, dtype=dtype),
                              start=torch.full((3,), 0.3, dtype=dtype), cutoff=1.0,
                              trainable=("table",))
    return lj, tab


@pytest.mark.parametrize("which", [0, 1])
def test_func_grad_gives_the_kernel_forces_and_strained_virial(fake_kernels, which):
    x, cell, pairs = _frame()
    model = _models()[which]
    strain = torch.tensor([[0.01, 0.002, 0.0], [0.0, -0.005, 0.001], [0.003, 0.0, 0.004]],
                          dtype=x.dtype)
    energy = lambda p, s: model._energy_wrapped(p, cell, pairs, s)
    (gx, gs), e = torch.func.grad_and_value(energy, argnums=(0, 1))(x, strain)
    want_e, want_f, want_w = _fake_evaluate(model, x, cell, pairs.index, strain)
    assert torch.equal(e, want_e)
    assert torch.equal(gx, -want_f)
    assert torch.equal(gs, -want_w)
    # an independent route: the oracle's own autograd of the strained energy
    ref = torch.func.grad(lambda s: _oracle_energy(model, x, s, cell, pairs.index))(strain)
    assert torch.allclose(gs, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("which", [0, 1])
def test_parameter_gradients_through_autograd(fake_kernels, which):
    x, cell, pairs = _frame()
    model = _models()[which]
    e = model._energy_wrapped(x, cell, pairs) * 3.0
    params = [p for p in model._differentiable_parameters() if p is not None]
    grads = torch.autograd.grad(e, params)
    want = _fake_parameter_gradient(model, x, cell, pairs.index, None) * 3.0
    if which == 0:
        assert torch.allclose(grads[0], want[0]) and torch.allclose(grads[1], want[1])
    else:
        assert torch.allclose(grads[0], want)


def test_second_derivative_raises(fake_kernels):
    x, cell, pairs = _frame()
    model = _models()[0]
    xr = x.clone().requires_grad_(True)
    e = model._energy_wrapped(xr, cell, pairs)
    (g,) = torch.autograd.grad(e, xr, create_graph=True)
    with pytest.raises(RuntimeError, match="second derivatives"):
        torch.autograd.grad(g.square().sum(), model.sigma)


def test_vmap_over_frames(fake_kernels):
    frames = [_frame(seed=s) for s in range(3)]
    x = torch.stack([f[0] for f in frames])
    cell, pairs = frames[0][1], frames[0][2]
    model = _models()[0]
    batched = torch.func.vmap(lambda p: model._energy_wrapped(p, cell, pairs))(x)
    one_by_one = torch.stack([model._energy_wrapped(p, cell, pairs) for p in x])
    assert torch.equal(batched, one_by_one)
    gb = torch.func.vmap(torch.func.grad(lambda p: model._energy_wrapped(p, cell, pairs)))(x)
    assert torch.equal(gb[1], torch.func.grad(lambda p: model._energy_wrapped(p, cell, pairs))(x[1]))


def test_quantity_generic_route_matches_the_oracle(fake_kernels):
    """A user energy (an LJ model plus a harmonic restraint on one atom) goes through
    torch.func as in the reference; the restraint's force appears on its atom only."""
    x, cell, pairs = _frame()
    model = _models()[0]
    anchor = x[0].clone() + 0.01

    def energy(positions, cell_, pairs_, strain=None):
        return model._energy_wrapped(positions, cell_, pairs_, strain) + \
            5.0 * (positions[0] - anchor).square().sum()

    e, f, w = quantity.energy_force_virial(energy)(x, cell, pairs)
    e0, f0, w0 = _fake_evaluate(model, x, cell, pairs.index, None)
    restraint = -10.0 * (x[0] - anchor)
    assert torch.allclose(e, e0 + 5.0 * (x[0] - anchor).square().sum())
    assert torch.allclose(f[0], f0[0] + restraint) and torch.allclose(f[1:], f0[1:])
    assert torch.allclose(w, w0)
