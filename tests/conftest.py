"""Shared test setup: repository root on sys.path and the `gpu` marker."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the C ABI on the GPU")


@pytest.fixture(scope="session")
def gpu_available():
    import v3dsfmtoolkit_amd as v

    n = v.device_count()
    if n <= 0:
        pytest.fail("no HIP device visible: the gpu tests need an MI355X")
    return n
