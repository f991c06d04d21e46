import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmcnla_amd on the GPU)")
    config.addinivalue_line("markers", "slow: larger sizes")


@pytest.fixture(scope="session")
def gpu():
    """Checks a GPU is present; GPU tests must run where one is (they never fall back)."""
    import mcnla_amd as mc
    n = mc.device_count()
    assert n >= 1, "no GPU visible: gpu-marked tests must run on a GPU machine"
    return n
