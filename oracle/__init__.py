"""ORACLE — test infrastructure, not product.

A CPU restatement of NablaDyn's neighbor-list-and-pair-force path, used only as the checker:
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it;
nothing under ``nabladyn_amd/`` does.

- ``oracle.neighbors``: the cell list (C restatement in ``neighbors.c`` via ctypes, exact in
  the list's dtype, reference column order) and an independent numpy brute force.
- ``oracle.potentials``: LJ and tabulated energy, forces, virial in fp64 (closed forms) and a
  torch-autograd restatement of the reference's ``_energy_wrapped`` + ``torch.func`` path.
- ``oracle.md``: velocity Verlet, the Nose-Hoover chain and ``maxwell_boltzmann``.

Every function cites the reference file:line it restates. Parity status per function is in
each module's header.
"""
