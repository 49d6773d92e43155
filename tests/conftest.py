import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liborrery_lbm.so")
    config.addinivalue_line("markers", "slow: long CPU run (excluded from the quick CPU suite by -m 'not slow')")


@pytest.fixture(scope="session")
def hip():
    """The HIP library, loaded; GPU tests fail (not skip) when it is missing."""
    from orrery_amd import abi
    return abi.load()
