"""``ndn_pair_efw_noalloc`` (the entry point that allocates nothing) against ``ndn_pair_efw``
(the same kernels with a caller's workspace), called through the C ABI.

Two calls in a row on one GPU, the second with a smaller input (fewer atoms, fewer columns, so
fewer blocks write partials into the shared scratch). Energy and virial are summed in a fixed
order, so they must match bit for bit; forces are added with atomics in any order, so they match
to 1e-6 of max|F| in fp32 and 1e-13 in fp64. LJ in fp32 and a two-type table in fp64.
"""

import ctypes

import numpy as np
import pytest
import torch

from nabladyn_amd import _native, potential, quantity, space
from nabladyn_amd.neighbors import NeighborList
from oracle import neighbors as on

pytestmark = pytest.mark.gpu


def _system(n, length, seed, dtype, device, kind):
    per = round(n ** (1 / 3))
    x = torch.tensor(on.jittered_lattice(per, length, 0.02, seed=seed), dtype=dtype, device=device)
    cell = torch.full((3,), length, dtype=dtype, device=device)
    if kind == "lj":
        model = potential.lennard_jones(
            atom_types=torch.zeros(n, dtype=torch.int64, device=device),
            sigma=torch.tensor([0.34], dtype=dtype, device=device),
            epsilon=torch.tensor([0.996], dtype=dtype, device=device), cutoff=1.0)
    else:
        rng = np.random.default_rng(seed)
        r = 0.3 + np.arange(200) * 0.9 / 199
        table = np.stack([4 * e * ((0.34 / r) ** 12 - (0.34 / r) ** 6) + rng.normal(0, 0.05, 200)
                          for e in (0.9, 0.7, 0.5)])
        table[:, -1] = 0.0
        model = potential.tabulated(
            atom_types=torch.tensor(rng.integers(0, 2, n), device=device),
            table=torch.tensor(table, dtype=dtype, device=device),
            start=torch.full((3,), 0.3, dtype=dtype, device=device), cutoff=1.2)
    nl = NeighborList(model.cutoff, skin=0.1)
    nl.allocate(x, cell)
    return space.wrap(x, cell)[0], cell, nl.tensors(), model


def _noalloc(x, cell, pairs, model):
    pot, keep = model._native()
    e = torch.empty((), dtype=x.dtype, device=x.device)
    f = torch.empty((x.shape[0], 3), dtype=x.dtype, device=x.device)
    w = torch.empty((3, 3), dtype=x.dtype, device=x.device)
    index = pairs.index.contiguous()
    _native.call("pair_efw_noalloc", _native.dtype_tag(x.dtype), _native.ptr(x),
                 _native.ptr(cell), ctypes.c_int64(x.shape[0]), _native.ptr(index),
                 ctypes.c_int64(index.shape[1]), ctypes.byref(pot), _native.ptr(e), _native.ptr(f),
                 _native.ptr(w), _native.stream(x.device))
    del keep
    return e, f, w


@pytest.mark.parametrize("dtype,kind,tol", [(torch.float32, "lj", 1e-6), (torch.float64, "table", 1e-13)])
def test_noalloc_equals_workspace_entry_point(device, dtype, kind, tol):
    for n, length, seed in ((4096, 5.8, 1), (512, 2.9, 2)):
        print(f"noalloc {kind} {dtype}: {n} atoms", flush=True)
        x, cell, pairs, model = _system(n, length, seed, dtype, device, kind)
        got = _noalloc(x, cell, pairs, model)
        # ndn_pair_efw, the same kernels with a workspace: the half list alone
        want = quantity._evaluate(model, x, cell, pairs._replace(full=None), energy=True,
                                  forces=True, virial=True)
        torch.cuda.synchronize()
        assert torch.equal(got[0], want[0]), (float(got[0]), float(want[0]))
        assert torch.equal(got[2], want[2])
        scale = float(want[1].abs().max())
        assert float((got[1] - want[1]).abs().max()) <= tol * scale
