"""Pins the pair-potential oracle: lj_frames.npz fp64 labels (2 types), the documented argon_512
numbers (examples/README.md:54-56), agreement of the closed-form and autograd routes, and the
invariants the Catmull-Rom table fixes (exact for quadratics vanishing at rc, tangent line below
the first knot, linear in the knots, C1 at knots, zero at the cutoff)."""

import numpy as np
import pytest
import torch

from oracle import neighbors as on
from oracle import potentials as op


def test_lj_frames_labels(lj_frames):
    d = lj_frames
    for f in range(d["positions"].shape[0]):
        xw, _ = on.wrap(d["positions"][f], d["cells"][f])
        b = on.build(xw, d["cells"][f], 1.0)
        e, forces, virial = op.lj_efw(xw, d["cells"][f], b.pairs, d["atom_types"], d["sigma"],
                                      d["epsilon"], 1.0)
        assert abs(e - d["energy"][f]) <= 1e-12 * abs(d["energy"][f])
        assert np.abs(forces - d["forces"][f]).max() <= 1e-12 * np.abs(d["forces"][f]).max()
        assert np.abs(virial - d["virial"][f]).max() <= 1e-12 * np.abs(d["virial"][f]).max()


def test_argon_documented_numbers(argon):
    xw, _ = on.wrap(argon["positions"], argon["cell"])
    b = on.build(xw, argon["cell"], 1.1)
    e, forces, virial = op.lj_efw(xw, argon["cell"], b.pairs, np.zeros(512, np.int64), [0.34],
                                  [0.996], 1.0)
    kinetic = 0.5 * (39.948 * argon["velocities"] ** 2).sum()
    pressure = (2 * kinetic + np.trace(virial)) / (3 * np.prod(argon["cell"])) / 0.0602214076
    assert f"{e:.2f}" == "-2526.17"
    assert f"{np.linalg.norm(forces, axis=1).max():.2f}" == "386.06"
    assert round(pressure) == 745


def _frame(n=400, length=3.4, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, length, (n, 3))
    cell = np.full(3, length)
    xw, _ = on.wrap(x, cell)
    return xw, cell, on.build(xw, cell, 1.3)


def test_lj_closed_form_equals_autograd():
    xw, cell, b = _frame()
    types = np.random.default_rng(2).integers(0, 2, len(xw))
    sigma, eps = np.array([0.34, 0.3, 0.32]), np.array([0.9, 1.1, 0.7])
    e, f, w = op.lj_efw(xw, cell, b.pairs, types, sigma, eps, 1.2)
    energy = lambda p, s, c, i: op.t_lj_energy(p, s, c, i, torch.tensor(types), torch.tensor(sigma),
                                               torch.tensor(eps), 1.2)
    e2, f2, w2 = op.torch_efw(energy, torch.tensor(xw), torch.tensor(cell), torch.tensor(b.pairs).long())
    assert abs(float(e2) - e) <= 1e-12 * abs(e)
    assert np.abs(f2.numpy() - f).max() <= 1e-11 * np.abs(f).max()
    assert np.abs(w2.numpy() - w).max() <= 1e-11 * np.abs(w).max()


def _table_of(function, knots=40, start=0.3, cutoff=1.2):
    r = start + np.arange(knots) * (cutoff - start) / (knots - 1)
    table = function(r)[None, :]
    table[:, -1] = 0.0
    return table, np.array([start]), (cutoff - start) / (knots - 1)


def test_table_exact_for_quadratics_and_tangent_below():
    a, b_ = 3.0, -2.0
    q = lambda r: a * (r - 1.2) ** 2 + b_ * (r - 1.2)
    dq = lambda r: 2 * a * (r - 1.2) + b_
    table, start, h = _table_of(q)
    r = np.linspace(0.3, 1.2, 1001)[:-1]
    rows = np.zeros_like(r, dtype=np.int64)
    u, du = op.hermite(r, rows, table, start, np.array([h]))
    assert np.abs(u - q(r)).max() < 1e-12
    assert np.abs(du - dq(r)).max() < 1e-10
    below = np.linspace(0.05, 0.3, 50)
    u, du = op.hermite(below, np.zeros_like(below, dtype=np.int64), table, start, np.array([h]))
    tangent = q(0.3) + dq(0.3) * (below - 0.3)
    assert np.abs(u - tangent).max() < 1e-12
    assert np.abs(du - dq(0.3)).max() < 1e-10


def test_table_linear_c1_zero_at_cutoff():
    rng = np.random.default_rng(4)
    t1 = rng.normal(size=(1, 30))
    t2 = rng.normal(size=(1, 30))
    t1[:, -1] = t2[:, -1] = 0
    start, h = np.array([0.3]), np.array([(1.2 - 0.3) / 29])
    r = np.linspace(0.1, 1.2, 777)
    rows = np.zeros_like(r, dtype=np.int64)
    u1, _ = op.hermite(r, rows, t1, start, h)
    u2, _ = op.hermite(r, rows, t2, start, h)
    u12, _ = op.hermite(r, rows, 2.0 * t1 - 0.5 * t2, start, h)
    assert np.abs(u12 - (2.0 * u1 - 0.5 * u2)).max() < 1e-12
    knots = 0.3 + np.arange(1, 29) * h[0]
    eps = 1e-10
    _, left = op.hermite(knots - eps, np.zeros(28, np.int64), t1, start, h)
    _, right = op.hermite(knots + eps, np.zeros(28, np.int64), t1, start, h)
    # C1: slopes agree across every knot (up to the curvature times 2 eps)
    assert np.abs(left - right).max() < 1e-5 * np.abs(left).max()
    u_end, _ = op.hermite(np.array([1.2]), np.zeros(1, np.int64), t1, start, h)
    assert abs(u_end[0]) < 1e-12


def test_table_closed_form_equals_autograd():
    xw, cell, b = _frame(seed=3)
    rng = np.random.default_rng(5)
    types = rng.integers(0, 2, len(xw))
    table = rng.normal(size=(3, 25))
    table[:, -1] = 0
    start = np.array([0.3, 0.25, 0.35])
    e, f, w = op.table_efw(xw, cell, b.pairs, types, table, start, 1.2)
    spacing = (1.2 - start) / 24
    energy = lambda p, s, c, i: op.t_table_energy(p, s, c, i, torch.tensor(types), torch.tensor(table),
                                                  torch.tensor(start), torch.tensor(spacing), 1.2)
    e2, f2, w2 = op.torch_efw(energy, torch.tensor(xw), torch.tensor(cell), torch.tensor(b.pairs).long())
    assert abs(float(e2) - e) <= 1e-11 * max(1.0, abs(e))
    assert np.abs(f2.numpy() - f).max() <= 1e-10 * np.abs(f).max()
    assert np.abs(w2.numpy() - w).max() <= 1e-10 * np.abs(w).max()


def test_table_energy_gradient_is_knot_weight():
    xw, cell, b = _frame(n=150, seed=6)
    types = np.zeros(len(xw), np.int64)
    table = np.random.default_rng(7).normal(size=(1, 12))
    table[:, -1] = 0
    start = np.array([0.3])
    grad = op.table_energy_gradient(xw, cell, b.pairs, types, table, start, 1.2)
    e = op.table_efw(xw, cell, b.pairs, types, table, start, 1.2)[0]
    assert abs((grad * table).sum() - e) < 1e-9 * max(1.0, abs(e))  # linear, no constant term
    assert grad[0, -1] == 0.0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_route_square_root_is_correctly_rounded(dtype):
    """The route's square root is the IEEE one the reference's CUDA platform computes: for each
    sampled x, no neighbor of the result in the dtype is closer to sqrt(x) (checked exactly in
    rationals: r is nearest iff (r - ulp/2)^2 <= x <= (r + ulp/2)^2). Its derivative is
    1 / (2 r)."""
    from fractions import Fraction

    rng = np.random.default_rng(0)
    x = torch.tensor(rng.uniform(0.01, 1.5, 3000), dtype=dtype)
    r = op.rounded_sqrt(x)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    for xv, rv in zip(x.numpy(), r.numpy()):
        up = Fraction(float(np.nextafter(rv, npdt(np.inf), dtype=npdt)))
        down = Fraction(float(np.nextafter(rv, npdt(0), dtype=npdt)))
        here = Fraction(float(rv))
        lo, hi = (down + here) / 2, (here + up) / 2
        assert lo * lo <= Fraction(float(xv)) <= hi * hi
    g = torch.func.grad(lambda v: op.rounded_sqrt(v).sum())(x)
    assert torch.equal(g, 1.0 / (2 * r))
