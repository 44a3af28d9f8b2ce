import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built native library")


@pytest.fixture(scope="session")
def golden():
    return dict(np.load(GOLDEN, allow_pickle=False))


@pytest.fixture(scope="session")
def oracle():
    from oracle_bindings import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle_bindings import Reference, reference_available
    if not reference_available():
        pytest.skip("reference build (oracle/_ref) not present")
    return Reference()


@pytest.fixture(scope="session")
def gpu():
    """One HIP context for the whole GPU session (tests run in one process)."""
    from anacapa_amd.integrator import HipPathIntegrator
    be = HipPathIntegrator(0)
    if not be.is_valid():
        pytest.fail(f"HIP backend not valid: {be.last_error()}")
    print(f"[gpu] device: {be.name()}", flush=True)
    yield be
    be.close()
