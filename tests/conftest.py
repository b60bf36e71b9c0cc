import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    """GPU tests fail (not skip) without a HIP device: they are the parity
    tests of the HIP path and must never pass vacuously."""
    from cuda_fot_amd import capi
    if capi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True
