"""Shared pytest configuration.

Marker ``gpu``: the test needs an MI355X and calls the HIP library through the C
ABI.  ``python -m pytest tests -m "not gpu"`` runs the CPU suite (oracle pinning,
host logic, library loading); ``-m gpu`` runs the parity tests on the GPU.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def handle():
    import high_precision_blas1_amd as H
    h = H.Handle()
    yield h
    h.close()


@pytest.fixture
def rng():
    return np.random.default_rng(12345)


def say(*a):
    """Print a case description before a GPU step, flushed (names the case if it dies)."""
    print(*a, flush=True)
