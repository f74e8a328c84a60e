"""Test configuration: the ``gpu`` marker, fixtures and shared generators.

``-m "not gpu"`` runs the oracle, host logic and library-loading tests on the CPU;
``-m gpu`` runs the parity tests of the HIP kernels against the oracle on an MI355X.
"""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests")


def pytest_collection_modifyitems(config, items):
    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason="no HIP device in this process")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def device():
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def argon():
    return dict(np.load(GOLDEN / "argon_512.npz"))


@pytest.fixture(scope="session")
def lj_frames():
    return dict(np.load(GOLDEN / "lj_frames_subset.npz"))
