import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI library)")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(autouse=True)
def _close_open_device():
    """The device library holds one state per process: a test that fails
    before closing its Device must not leave it open for the next test."""
    yield
    from lare3d_amd.device import Device
    if Device._open is not None:
        Device._open.close()
