import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liblyap.so on cuda:0)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    from chaos_oneapi_amd import lyap as L

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ctx = L.Context(0)
    yield ctx
    ctx.close()
