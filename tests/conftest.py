import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950) and the HIP engine")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One device context for the GPU tests of the session.  A GPU test run
    on a machine without a device fails here (it is selected with -m gpu)."""
    from tight_binding_amd import kpm

    n = kpm.device_count()
    if n == 0:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    ctx = kpm.Context(0)
    yield ctx
    ctx.close()
