"""Shared test setup: import paths, the `gpu` marker and backend helpers.

Tests marked `gpu` need a gfx950 device and the built libdrape_hip.so; everything else runs on the CPU (the FP64
oracle and the host logic).
"""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "drape-2_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def hip():
    import drape_amd
    if not drape_amd.hip_available():
        pytest.fail("gpu test run without a usable gfx950 device / libdrape_hip.so")
    return drape_amd


def max_distance(a, b) -> float:
    return float(np.max(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64), axis=1)))
