"""Synthetic inputs of the benchmark configurations, generated from fixed seeds.

- ``argon_lattice``: a jittered simple-cubic lattice of argon at 21 nm^-3 (configs 1 and 2).
- ``ragged_systems``: 256 independent systems with log-uniform sizes, orthorhombic boxes of
  aspect ratios in [1, 2], half LJ and half a 200-knot table (config 3).
- ``cg_frames``: coarse-grained frames with LJ target forces plus noise (config 4).
- ``request_stream``: a Poisson stream of energy/force and short-MD requests (config 5).

Positions are generated in fp64 and cast last. Sites sit at (k + 1/2) a; jitter is uniform.
"""

import math

import numpy as np
import torch

ARGON_MASS = 39.948  # Da
ARGON_SIGMA = 0.34  # nm
ARGON_EPSILON = 0.996  # kJ/mol
DENSITY = 21.0  # nm^-3


def lattice(counts, lengths, jitter: float, rng: np.random.Generator, n: int | None = None):
    """Sites of a counts[0] x counts[1] x counts[2] lattice filling a box of `lengths`,
    x fastest, the first `n` of them, each jittered uniformly by +-jitter."""
    counts = [int(c) for c in counts]
    a = np.asarray(lengths, dtype=np.float64) / np.asarray(counts)
    gz, gy, gx = np.meshgrid(*(np.arange(c, dtype=np.float64) for c in counts[::-1]), indexing="ij")
    sites = (np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1) + 0.5) * a
    if n is not None:
        sites = sites[:n]
    return sites + rng.uniform(-jitter, jitter, size=sites.shape)


def argon_lattice(per_axis: int, *, density: float = DENSITY, jitter: float = 0.02, seed: int = 0,
                  length: float | None = None):
    """A cubic box of per_axis^3 argon atoms at `density` (or of edge `length`):
    (positions [N, 3] fp64, cell [3] fp64)."""
    n = per_axis**3
    if length is None:
        length = (n / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    x = lattice((per_axis,) * 3, (length,) * 3, jitter, rng)
    return x, np.full(3, length)


def lj_table(knots: int = 200, start: float = 0.3, cutoff: float = 1.2, noise: float = 0.05,
             seed: int = 3, sigma: float = ARGON_SIGMA, epsilon: float = ARGON_EPSILON):
    """A table of `knots` values on [start, cutoff]: LJ plus N(0, noise), last knot 0."""
    rng = np.random.default_rng(seed)
    r = start + np.arange(knots) * (cutoff - start) / (knots - 1)
    u = 4 * epsilon * ((sigma / r) ** 12 - (sigma / r) ** 6) + rng.normal(0, noise, knots)
    u[-1] = 0.0
    return u


def ragged_systems(count: int = 256, *, low: int = 64, high: int = 32768, seed: int = 2,
                   density: float = DENSITY, min_length: float | None = None):
    """Config 3's systems: a list of dicts with positions [n, 3], cell [3], kind ('lj'/'table').

    Sizes are log-uniform in [low, high]; aspect ratios (1, r1, r2) with r uniform in [1, 2];
    the volume n / density fixes the box. With ``min_length`` every edge is at least that long
    (the box grows and the density drops for the few small systems that need it).
    """
    rng = np.random.default_rng(seed)
    systems = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(low), math.log(high)))))
        ratios = np.array([1.0, rng.uniform(1, 2), rng.uniform(1, 2)])
        volume = n / density
        scale = (volume / ratios.prod()) ** (1 / 3)
        lengths = ratios * scale
        if min_length is not None and lengths.min() < min_length:
            lengths = lengths * (min_length / lengths.min())
        per = np.maximum(1, np.round(lengths / (lengths.prod() / n) ** (1 / 3))).astype(int)
        while per.prod() < n:
            per[np.argmin(per / lengths)] += 1
        x = lattice(per, lengths, 0.02, rng, n=n)
        systems.append({"positions": x, "cell": lengths, "kind": "lj" if k % 2 == 0 else "table"})
    return systems


def cg_frames(frames: int = 64, beads: int = 4096, *, density: float = 8.0, seed: int = 4,
              sigma: float = 0.47, epsilon: float = 3.5, force_noise: float = 5.0):
    """Config 4's frames: positions [F, N, 3] (jittered 16^3 lattices), cells [F, 3], and the
    LJ(sigma, epsilon) parameters whose forces (plus noise) are the targets."""
    rng = np.random.default_rng(seed)
    per = round(beads ** (1 / 3))
    length = (beads / density) ** (1 / 3)
    xs = np.stack([lattice((per,) * 3, (length,) * 3, 0.05, rng) for _ in range(frames)])
    cells = np.full((frames, 3), length)
    return xs, cells, {"sigma": sigma, "epsilon": epsilon, "force_noise": force_noise, "rng": rng}


def request_stream(count: int, qps: float, *, seed: int = 4, low: int = 256, high: int = 262144,
                   md_fraction: float = 0.3, md_steps: int = 500):
    """Config 5's requests: arrival time (s), kind ('efw' or 'md'), size, potential, steps."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / qps, count))
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(low), math.log(high)))))
        kind = "md" if rng.random() < md_fraction else "efw"
        out.append({"arrival": float(t[k]), "kind": kind, "atoms": n,
                    "potential": "lj" if rng.random() < 0.5 else "table",
                    "steps": md_steps if kind == "md" else 0, "seed": int(rng.integers(1 << 30))})
    return out


def to_device(array, dtype, device):
    return torch.tensor(np.asarray(array), dtype=dtype, device=device)
