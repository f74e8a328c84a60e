"""Pins the neighbor-list oracle: the C restatement against an independent numpy brute force,
the counts the reference documents for argon_512 (README.md:166-172), closed forms of the reuse
rule (update/update.cuh:27-116), and the layout invariants of the pair index."""

import math

import numpy as np
import pytest

from oracle import neighbors as on


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_argon_documented_counts(argon, dtype):
    xw, _ = on.wrap(argon["positions"].astype(dtype), argon["cell"].astype(dtype))
    cell = argon["cell"].astype(dtype)
    b = on.build(xw, cell, 1.1)
    assert b.pair_total == 29014
    assert b.capacity == 36268
    assert b.all_pairs
    assert on.build(xw, cell, 1.0).pair_total == 21398
    np.testing.assert_array_equal(on.pair_set(b.pairs, b.pair_total), on.brute_pairs(xw, cell, 1.1))


CASES = [
    (3000, (6.1, 5.3, 7.7), 1.2),   # 27-bin path
    (2500, (2.5, 6.0, 6.0), 1.2),   # a short axis
    (1500, (9.0, 2.3, 2.4), 1.1),   # two short axes
    (700, (3.0, 3.0, 3.0), 1.2),    # all-pairs path
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_c_restatement_equals_brute_force(case, dtype):
    n, cell, radius = case
    cell = np.array(cell, dtype=dtype)
    rng = np.random.default_rng(n)
    x = (rng.uniform(0, 1, (n, 3)) * cell + rng.normal(0, 2, (n, 3))).astype(dtype)
    xw, crossings = on.wrap(x, cell)
    assert (xw >= 0).all() and (xw < cell).all()
    np.testing.assert_allclose(xw + crossings * cell, x, rtol=0, atol=1e-4 if dtype == np.float32 else 1e-12)
    b = on.build(xw, cell, radius)
    brute = on.brute_pairs(xw, cell, radius)
    np.testing.assert_array_equal(on.pair_set(b.pairs, b.pair_total), brute)
    # each unordered pair once; padding (k mod N, k mod N)
    assert len(np.unique(brute, axis=0)) == b.pair_total
    k = np.arange(b.pair_total, b.capacity)
    np.testing.assert_array_equal(b.pairs[0, b.pair_total:], k % n)
    np.testing.assert_array_equal(b.pairs[1, b.pair_total:], k % n)
    if b.all_pairs:
        first, second = b.pairs[:, : b.pair_total]
        assert (first < second).all()
        order = np.lexsort((second, first))
        np.testing.assert_array_equal(order, np.arange(b.pair_total))
    else:
        # columns grouped by center in (bin, id) order: the first row changes only at run ends
        first = b.pairs[0, : b.pair_total]
        runs = first[np.r_[True, first[1:] != first[:-1]]]
        assert len(runs) == len(np.unique(runs))


def test_boundary_pairs_follow_unfused_arithmetic():
    """Pairs whose r^2 lies within a few ulp of R^2: the oracle's decision is the unfused
    (dx dx + dy dy) + dz dz, and a fused evaluation would decide some of them differently."""
    rng = np.random.default_rng(0)
    radius = np.float32(1.2)
    r2 = radius * radius
    cell = np.full(3, 9.0, dtype=np.float32)
    centers = rng.uniform(3, 6, (4000, 3)).astype(np.float32)
    direction = rng.normal(size=(4000, 3))
    direction /= np.linalg.norm(direction, axis=1, keepdims=True)
    scale = 1.0 + rng.uniform(-3e-7, 3e-7, 4000)
    others = (centers + direction * 1.2 * scale[:, None]).astype(np.float32)
    d = (centers - others).astype(np.float32)
    unfused = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    fused = np.float32(np.float64(d[:, 2]) * d[:, 2] + np.float64(np.float32(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])))
    assert (unfused < r2).any() and (unfused >= r2).any()
    assert ((unfused < r2) != (fused < r2)).any()  # the rule matters
    for k in range(0, 4000, 500):
        x = np.stack([centers[k], others[k]])
        b = on.build(x, cell, float(radius))
        assert b.pair_total == int(unfused[k] < r2)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reuse_threshold_closed_form(dtype):
    eps = np.finfo(dtype).eps
    cell = np.array([4.0, 5.0, 6.0], dtype=dtype)
    cutoff, radius = 1.0, 1.2
    rebuild, status, dims, widths, reuse = on.prepare_update(cell, cell, cutoff, radius, [3, 4, 5])
    allowance = dtype(radius) * dtype(1) - dtype(cutoff) - dtype(32) * dtype(eps) * dtype(6.0)
    assert rebuild == 0 and status == 0
    assert reuse[0] == allowance * allowance / dtype(4)
    assert math.isclose(math.sqrt(reuse[0]), 0.1, rel_tol=1e-3)  # half the skin, less the guard
    np.testing.assert_array_equal(reuse[1:], np.ones(3, dtype=dtype))
    # forced rebuild: an infinite reference cell
    assert on.prepare_update(cell, np.full(3, np.inf, dtype=dtype), cutoff, radius, [3, 4, 5])[0] == 1
    # a compressed cell rebuilds when R s - rc - 32 eps L <= 0
    assert on.prepare_update(cell * dtype(0.8), cell, cutoff, radius, [3, 4, 5])[0] == 1
    # invalid cell
    assert on.prepare_update(np.array([1.9, 5, 6], dtype=dtype), cell, cutoff, radius, [3, 4, 5])[1] == 1
    assert on.prepare_update(np.array([np.nan, 5, 6], dtype=dtype), cell, cutoff, radius, [3, 4, 5])[1] == 1


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_reuse_check_decision(dtype):
    cell = np.full(3, 5.0, dtype=dtype)
    rng = np.random.default_rng(1)
    x0 = rng.uniform(0, 5, (100, 3)).astype(dtype)
    rebuild, status, _, _, reuse = on.prepare_update(cell, cell, 1.0, 1.2, [4, 4, 4])
    threshold = math.sqrt(reuse[0])
    moved = x0.copy()
    moved[17, 0] = x0[17, 0] + dtype(0.99 * threshold)
    assert on.check_reuse(on.wrap(moved, cell)[0], x0, cell, reuse, rebuild) == (0, 0)
    moved[17, 0] = x0[17, 0] + dtype(1.01 * threshold)
    assert on.check_reuse(on.wrap(moved, cell)[0], x0, cell, reuse, rebuild) == (1, 0)
    # across the periodic face: the minimum image keeps a tiny move tiny
    x0[3] = [dtype(4.99999), 1, 1]
    moved = x0.copy()
    moved[3, 0] = dtype(0.00001)
    assert on.check_reuse(moved, x0, cell, reuse, rebuild)[0] == 0
    moved[5, 2] = np.nan
    assert on.check_reuse(moved, x0, cell, reuse, rebuild)[1] == on.NON_FINITE
