import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the dmrgx C ABI)")
    config.addinivalue_line("markers", "slow: long CPU-side oracle runs")


@pytest.fixture(scope="session")
def gpu():
    """Initialise the HIP library once; the GPU tests fail loudly without it."""
    import dmrg_amd
    dmrg_amd.init(0)
    return dmrg_amd
