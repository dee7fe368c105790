import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle.bind import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reflib():
    from oracle.bind import RefLib
    if not RefLib.available():
        pytest.skip("reference build oracle/_ref not present")
    return RefLib()


@pytest.fixture(scope="session")
def ctx():
    """HIP context on device 0.  No skip: a GPU test without a device fails."""
    from poisson_caustic_design_amd import capi
    return capi.Context(0)


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
