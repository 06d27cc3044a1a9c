"""Shared fixtures: import paths for the engine's Python mirror and the oracle.

GPU tests carry @pytest.mark.gpu and call the HIP kernels through the C ABI
(libgbo.so via gbo.py); CPU tests check the oracle, host logic and the
library's exported symbols.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "graph-bit-ops_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gbo():
    import gbo as m
    return m


@pytest.fixture(scope="session")
def po():
    import pyoracle as m
    return m


def bernoulli_x(n, p, seed):
    rng = np.random.default_rng(seed)
    return (rng.random(n) < p).astype(np.float32)
