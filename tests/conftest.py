import os
import sys

import pytest

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpsl_amd.so")


@pytest.fixture(scope="session")
def psl():
    """The HIP library; GPU tests fail loudly when it or the device is missing."""
    import periodicshiftlbm_amd as P
    P.lib()
    P.require_device()
    return P


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle
