"""Shared test setup: import paths, the `gpu` marker, session fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "cuda-sort_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def cs():
    import cusort_hip
    cusort_hip.lib()
    return cusort_hip


@pytest.fixture(scope="session")
def gpu(cs):
    import hiprt
    assert hiprt.device_count() > 0, "GPU tests need a HIP device"
    hiprt.set_device(0)
    return hiprt
