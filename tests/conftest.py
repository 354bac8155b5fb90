"""Shared test set-up: import paths, the ``gpu`` marker, the reference's fixed seed."""

import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
PKG = REPO / "blackhole-nebula-simulator_amd"
for p in (str(REPO), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)

SEED = 20261014  # the reference's test seed (tests/conftest.py:7)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture
def rng() -> np.random.Generator:
    return np.random.default_rng(SEED)


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False
