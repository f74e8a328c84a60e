"""Supervised fitting of a potential to reference energies, forces and virials
(``training/__init__.py:1-9``).

A ``Dataset`` holds the labeled frames; ``fit`` trains the parameters and returns a ``Report``.
For LJ and tabulated models the gradient of the force and virial terms comes from the
library's parameter VJP kernel rather than a nested autograd graph.
"""

from ._dataset import Dataset
from ._fit import Report, batches, fit, pair_stack, width
from ._loss import CHANNELS, loss, predict, scales

__all__ = ("CHANNELS", "Dataset", "Report", "batches", "fit", "loss", "pair_stack", "predict",
           "scales", "width")
