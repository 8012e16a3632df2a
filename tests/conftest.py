import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libsmx.so")


@pytest.fixture(scope="session")
def canon_params():
    """A canonical parameter set (theta of the CPU-runnable configuration)."""
    from schwarmax_amd import synth
    from oracle import model as OM
    theta = synth.base_theta()
    ph, pb = OM.theta_to_params(theta)
    return theta, ph, pb


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
