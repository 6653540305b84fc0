import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU check")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def i2():
    import integrator2_amd
    integrator2_amd.lib()
    return integrator2_amd


@pytest.fixture(scope="session")
def gpu(i2):
    n = i2.device_count()
    if n < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return i2
