import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu_ctx():
    import geogram_amd as ga

    ctx = ga.Context(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def emu_ctx():
    import geogram_amd as ga

    ctx = ga.Context.emulated()
    yield ctx
    ctx.close()
