"""Shared parity helpers for the table tests (imported by tests, collected by none).

The bar for fp32 forces and virials: within 1e-5 relative of the reference, with
each result's error against an fp64 evaluation stated beside it. "The reference" for a table is
its own fp32 route, restated by the oracle: ``Tabulated._energy_wrapped``
(``potential/_tabulated/_potential.py:146-162``) differentiated by ``torch.func`` as
``quantity.energy_force_virial`` does (``quantity.py:93-117``), run in fp32 on the CPU.
"""

import numpy as np
import torch

from oracle import potentials as op

# max|dX| / max|X_route| for F and W, |dE| / max(1, |E_route|)
ROUTE_TOL = 1e-5


def rel(a, b) -> float:
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def table_route(x_wrapped, cell, pairs, atom_types, table, start, spacing, cutoff):
    """The reference's route in the dtype of ``x_wrapped``: (E, F [N, 3], W [3, 3]) as numpy."""
    dtype = torch.float32 if np.asarray(x_wrapped).dtype == np.float32 else torch.float64
    as_t = lambda a: torch.as_tensor(np.asarray(a)).to(dtype)
    index = torch.as_tensor(np.asarray(pairs)).long()
    types = torch.as_tensor(np.asarray(atom_types)).long()
    e, f, w = op.torch_efw(op.t_table_energy, as_t(x_wrapped), as_t(cell), index, types,
                           as_t(table), as_t(start), as_t(spacing), float(cutoff))
    return float(e), f.numpy(), w.numpy()


def lj_route(x_wrapped, cell, pairs, atom_types, sigma, epsilon, cutoff):
    """``LennardJones._energy_wrapped`` (``potential/_lennard_jones/_potential.py:209-234``)
    through the same route, in the dtype of ``x_wrapped``."""
    dtype = torch.float32 if np.asarray(x_wrapped).dtype == np.float32 else torch.float64
    as_t = lambda a: torch.as_tensor(np.asarray(a)).to(dtype)
    index = torch.as_tensor(np.asarray(pairs)).long()
    types = torch.as_tensor(np.asarray(atom_types)).long()
    e, f, w = op.torch_efw(op.t_lj_energy, as_t(x_wrapped), as_t(cell), index, types,
                           as_t(sigma), as_t(epsilon), float(cutoff))
    return float(e), f.numpy(), w.numpy()


def check_against_route(name, got, route, fp64, tol=ROUTE_TOL):
    """Asserts got (E, F, W) within `tol` of the fp32 route; prints both results' fp64 errors.
    Returns the errors as a dict."""
    e, f, w = got
    er, fr, wr = route
    e64, f64, w64 = fp64
    errors = {
        "F_vs_route": rel(f, fr), "W_vs_route": rel(w, wr),
        "E_vs_route": abs(e - er) / max(1.0, abs(er)),
        "F_vs_fp64": rel(f, f64), "W_vs_fp64": rel(w, w64),
        "E_vs_fp64": abs(e - e64) / max(1.0, abs(e64)),
        "route_F_vs_fp64": rel(fr, f64), "route_W_vs_fp64": rel(wr, w64),
    }
    print(f"{name}: " + ", ".join(f"{k} {v:.2e}" for k, v in errors.items()))
    assert errors["F_vs_route"] <= tol, errors
    assert errors["W_vs_route"] <= tol, errors
    assert errors["E_vs_route"] <= tol, errors
    return errors
