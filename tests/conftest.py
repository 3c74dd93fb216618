import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def gpu_ctx():
    from lblrtm_amd import lib
    if lib.load().lblx_device_count() < 1:
        pytest.fail("no HIP device visible")
    ctx = lib.Context(0)
    yield ctx
    ctx.close()
