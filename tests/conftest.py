import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    from tensorray_amd import _capi
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    ctx = _capi.Context(0)
    yield ctx
    ctx.close()
