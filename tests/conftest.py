import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on cuda:0)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ilm():
    import illuminaprt_amd as ilm
    ilm.load()
    return ilm


@pytest.fixture(scope="session")
def gpu_ctx():
    import illuminaprt_amd as ilm
    ilm.load()
    if not ilm.gpu_available():
        pytest.fail("test marked gpu but no HIP device is visible")
    ctx = ilm.Context(0)
    yield ctx
    ctx.close()
