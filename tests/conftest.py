import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbgpu.so on cuda:0)")


@pytest.fixture(scope="session")
def gpu_ctx():
    from bling_amd.runtime import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
