"""ORACLE — test infrastructure, not product: pair energies, forces and virials on the CPU.

Two restatements of the reference's pair potentials:

- ``lj_efw``/``table_efw``: closed-form fp64 numpy. E sums u over listed pairs with
  i != j and r^2 < rc^2; F_i = -dE/dx_i; W = sum over pairs of r_ij (x) F_ij with
  r_ij = x_i - x_j and F_ij the force on i from j, which equals -dE/dstrain at zero strain for a
  pair potential (``quantity.py:93-117``, ``space.py:51-64``).
  Reference: ``potential/_lennard_jones/_potential.py:113-151,209-234`` (shifted 12-6 per
  unordered type pair, ``pair_index`` rows), ``potential/_tabulated/_potential.py:59-91,146-162``
  (Catmull-Rom with virtual end knots, last knot 0, tangent below the first knot),
  ``space.py:11-19`` (minimum image with the near-face coordinate shifted first).
- ``torch_efw``: the reference's eager torch computation restated (gathers, ``where`` masks,
  ``torch.func.grad_and_value`` over positions and strain), in any dtype, on the CPU. It is the
  "reference autograd" the hand-written kernels and backward are compared with, and the CPU
  baseline's force path. Its square root is correctly rounded, as on the reference's CUDA
  platform (torch's CPU ``sqrt`` is not; see ``_RoundedSqrt``).

Parity status: pinned. ``lj_efw`` reproduces the fp64 labels of ``lj_frames.npz`` (2 types) to
~1e-13 relative and the documented argon_512 numbers (E = -2526.17 kJ/mol, max |F| = 386.06,
P = 745 bar). ``table_efw`` is pinned by invariants the Catmull-Rom form fixes: exact for any
quadratic vanishing at rc (including the tangent extension), linear in the knots, C1 at knots,
and by agreement with ``torch_efw`` (independent derivative route: autograd vs closed form).
"""

import numpy as np
import torch


def pair_index(first, second, num_types: int):
    """``_lennard_jones/_potential.py:143-151``: row a(2T - a - 1)/2 + b of the pair (a <= b)."""
    low = np.minimum(first, second)
    high = np.maximum(first, second)
    return low * (2 * num_types - low - 1) // 2 + high


def num_types_from_rows(rows: int) -> int:
    t = int(np.floor(np.sqrt(2 * rows)))
    assert t * (t + 1) // 2 == rows
    return t


def displacement(left, right, cell):
    """``space.py:11-19``: (left - left_shift) - (right - right_shift), elementwise in the dtype."""
    raw = left - right
    half = 0.5 * cell
    left_shift = np.where(raw > half, cell, 0.0).astype(left.dtype)
    right_shift = np.where(raw < -half, cell, 0.0).astype(left.dtype)
    return (left - left_shift) - (right - right_shift)


def _pairs_geometry(x, cell, pairs, cutoff):
    x = np.asarray(x, dtype=np.float64)
    cell = np.asarray(cell, dtype=np.float64)
    first = np.asarray(pairs[0], dtype=np.int64)
    second = np.asarray(pairs[1], dtype=np.int64)
    r = displacement(x[first], x[second], cell)
    r2 = (r * r).sum(axis=1)
    valid = (first != second) & (r2 < cutoff * cutoff)
    return first[valid], second[valid], r[valid], r2[valid]


def _accumulate(x, first, second, r, fscalar, energies):
    """F_ij = fscalar * r_ij; returns (E, F [N,3], W [3,3]) in fp64."""
    fij = fscalar[:, None] * r
    forces = np.zeros((x.shape[0], 3), dtype=np.float64)
    np.add.at(forces, first, fij)
    np.add.at(forces, second, -fij)
    virial = r.T @ fij
    return float(energies.sum()), forces, virial


def lj_efw(x, cell, pairs, atom_types, sigma, epsilon, cutoff):
    """Shifted 12-6 LJ in fp64: (E, F [N, 3], W [3, 3]); x already wrapped into [0, L)."""
    sigma = np.asarray(sigma, dtype=np.float64)
    epsilon = np.asarray(epsilon, dtype=np.float64)
    t = num_types_from_rows(sigma.size)
    first, second, r, r2 = _pairs_geometry(x, cell, pairs, cutoff)
    types = np.asarray(atom_types, dtype=np.int64)
    rows = pair_index(types[first], types[second], t) if t > 1 else np.zeros(len(first), np.int64)
    s2 = sigma[rows] ** 2
    e = epsilon[rows]
    p6 = (s2 / r2) ** 3
    pc = (s2 / (cutoff * cutoff)) ** 3
    energies = 4.0 * e * (p6 * p6 - p6) - 4.0 * e * (pc * pc - pc)
    # -du/dr / r = 24 eps (2 p6^2 - p6) / r^2
    fscalar = 24.0 * e * (2.0 * p6 * p6 - p6) / r2
    return _accumulate(np.asarray(x), first, second, r, fscalar, energies)


def padded_table(table):
    """``_tabulated/_potential.py:79-82``: [P, K + 2] knots with the last knot read as 0."""
    table = np.asarray(table, dtype=np.float64)
    values = np.concatenate([table[:, :-1], np.zeros_like(table[:, :1])], axis=1)
    before = 3 * values[:, :1] - 3 * values[:, 1:2] + values[:, 2:3]
    after = 3 * values[:, -1:] - 3 * values[:, -2:-1] + values[:, -3:-2]
    return np.concatenate([before, values, after], axis=1)


def hermite(r, rows, table, start, spacing):
    """u(r) and du/dr in fp64 (``_tabulated/_potential.py:59-91``)."""
    padded = padded_table(table)
    k = padded.shape[1] - 2
    start = np.asarray(start, dtype=np.float64)
    spacing = np.asarray(spacing, dtype=np.float64)
    h = spacing[rows]
    xk = (r - start[rows]) / h
    seg = np.clip(np.floor(xk), 0, k - 2).astype(np.int64)
    t = xk - seg
    u0, u1, u2, u3 = (padded[rows, seg + m] for m in range(4))
    s = np.maximum(t, 0.0)
    a = u2 - u0
    b = 2 * u0 - 5 * u1 + 4 * u2 - u3
    c = 3 * (u1 - u2) + u3 - u0
    cubic = u1 + 0.5 * s * (a + s * (b + s * c))
    tangent = np.where(t < 0, 0.5 * a * t, 0.0)
    du_dt = 0.5 * (a + 2 * s * b + 3 * s * s * c)
    du_dt = np.where(t < 0, 0.5 * a, du_dt)
    return cubic + tangent, du_dt / h


def table_efw(x, cell, pairs, atom_types, table, start, cutoff, spacing=None):
    """Tabulated pair potential in fp64: (E, F [N, 3], W [3, 3]); x already wrapped."""
    table = np.asarray(table, dtype=np.float64)
    start = np.asarray(start, dtype=np.float64)
    if spacing is None:
        spacing = (cutoff - start) / (table.shape[1] - 1)
    t = num_types_from_rows(table.shape[0])
    first, second, r, r2 = _pairs_geometry(x, cell, pairs, cutoff)
    types = np.asarray(atom_types, dtype=np.int64)
    rows = pair_index(types[first], types[second], t)
    dist = np.sqrt(r2)
    u, du = hermite(dist, rows, table, start, spacing)
    fscalar = -du / dist
    return _accumulate(np.asarray(x), first, second, r, fscalar, u)


def table_energy_gradient(x, cell, pairs, atom_types, table, start, cutoff):
    """dE/dtable [P, K] in fp64: E is linear in the knots, so each knot's weight summed."""
    table = np.asarray(table, dtype=np.float64)
    p, k = table.shape
    grad = np.zeros_like(table)
    base = np.zeros_like(table)
    e0 = table_efw(x, cell, pairs, atom_types, base, start, cutoff)[0]
    for row in range(p):
        for col in range(k - 1):
            unit = base.copy()
            unit[row, col] = 1.0
            grad[row, col] = table_efw(x, cell, pairs, atom_types, unit, start, cutoff)[0] - e0
    return grad


# ---------------------------------------------------------------------------------------------
# torch restatement of the reference's eager path (any dtype, CPU)
# ---------------------------------------------------------------------------------------------


def _t_displacement(left, right, cell):
    """``space.py:11-19`` in torch."""
    raw = left - right
    half_cell = 0.5 * cell
    left_shift = torch.where(raw > half_cell, cell, 0.0)
    right_shift = torch.where(raw < -half_cell, cell, 0.0)
    return (left - left_shift) - (right - right_shift)


def _t_strained(vectors, strain):
    """``space.py:51-64``."""
    return (
        vectors
        + vectors[..., 0:1] * strain[:, 0]
        + vectors[..., 1:2] * strain[:, 1]
        + vectors[..., 2:3] * strain[:, 2]
    )


def _t_pair_index(first, second, num_types):
    low, high = torch.minimum(first, second), torch.maximum(first, second)
    return low * (2 * num_types - low - 1) // 2 + high


def t_lj_energy(positions, strain, cell, index, atom_types, sigma, epsilon, cutoff):
    """``LennardJones._energy_wrapped`` (``_lennard_jones/_potential.py:209-234``) restated."""
    first, second = index[0], index[1]
    vectors = _t_displacement(positions[first], positions[second], cell)
    squared = _t_strained(vectors, strain).square().sum(dim=-1)
    valid = (first != second) & (squared < cutoff**2)
    safe = torch.where(valid, squared, 1.0)
    num_types = int(round((np.sqrt(8 * sigma.numel() + 1) - 1) / 2))

    def pair_energy(sq, s, e):
        power6 = (s.square() / sq).pow(3)
        return 4.0 * e * (power6.square() - power6)

    cutoff_energy = pair_energy(cutoff**2, sigma, epsilon)
    if num_types == 1:
        s, e, shift = sigma[0], epsilon[0], cutoff_energy[0]
    else:
        rows = _t_pair_index(atom_types[first], atom_types[second], num_types)
        s, e, shift = sigma[rows], epsilon[rows], cutoff_energy[rows]
    energy = pair_energy(safe, s, e) - shift
    return torch.where(valid, energy, 0.0).sum()


class _RoundedSqrt(torch.autograd.Function):
    """The square root rounded once to the dtype, as the reference's platform computes it
    (``torch.sqrt`` on CUDA is ``sqrtf``/``sqrt``, IEEE correctly rounded). torch's CPU routine is
    not: on an AVX-512 host about 0.7% of its fp32 and fp64 roots are 1 ulp off, which moves the
    table's knot coordinate (r - start) / h by an ulp. numpy's ``sqrt`` is the hardware's
    correctly rounded instruction. The derivative is torch's own, grad / (2 sqrt(x)) in the
    dtype."""

    @staticmethod
    def forward(x):
        return torch.from_numpy(np.sqrt(x.detach().cpu().numpy())).to(x.device)

    @staticmethod
    def setup_context(ctx, inputs, output):
        ctx.save_for_backward(output)

    @staticmethod
    def backward(ctx, grad):
        (root,) = ctx.saved_tensors
        return grad / (2 * root)


def rounded_sqrt(x):
    """``x.sqrt()`` correctly rounded (see ``_RoundedSqrt``)."""
    return _RoundedSqrt.apply(x)


def t_hermite(r, rows, table, start, spacing):
    """``_hermite_energy`` (``_tabulated/_potential.py:59-91``) restated."""
    num_knots = table.shape[1]
    values = torch.cat([table[:, :-1], torch.zeros_like(table[:, :1])], dim=1)
    before = 3 * values[:, :1] - 3 * values[:, 1:2] + values[:, 2:3]
    after = 3 * values[:, -1:] - 3 * values[:, -2:-1] + values[:, -3:-2]
    padded = torch.cat([before, values, after], dim=1).reshape(-1)
    x = (r - start[rows]) / spacing[rows]
    segment = x.floor().clamp(0, num_knots - 2)
    t = x - segment
    first = rows * (num_knots + 2) + segment.long()
    u0, u1, u2, u3 = (padded[first + k] for k in range(4))
    s = t.clamp(min=0)
    cubic = u1 + 0.5 * s * (
        (u2 - u0) + s * ((2 * u0 - 5 * u1 + 4 * u2 - u3) + s * (3 * (u1 - u2) + u3 - u0))
    )
    return cubic + torch.where(t < 0, 0.5 * (u2 - u0) * t, 0.0)


def t_table_energy(positions, strain, cell, index, atom_types, table, start, spacing, cutoff):
    """``Tabulated._energy_wrapped`` (``_tabulated/_potential.py:146-162``) restated."""
    first, second = index[0], index[1]
    vectors = _t_displacement(positions[first], positions[second], cell)
    squared = _t_strained(vectors, strain).square().sum(dim=-1)
    valid = (first != second) & (squared < cutoff**2)
    r = rounded_sqrt(torch.where(valid, squared, cutoff**2))
    num_types = int(round((np.sqrt(8 * table.shape[0] + 1) - 1) / 2))
    rows = _t_pair_index(atom_types[first], atom_types[second], num_types)
    energy = t_hermite(r, rows, table, start, spacing)
    return torch.where(valid, energy, 0.0).sum()


def torch_efw(energy, positions, cell, *args):
    """``quantity.energy_force_virial`` (``quantity.py:93-117``): grad_and_value over (x, strain)."""
    grad_and_value = torch.func.grad_and_value(energy, argnums=(0, 1))
    strain = positions.new_zeros((3, 3))
    (gx, gs), e = grad_and_value(positions, strain, cell, *args)
    return e, -gx, -gs
