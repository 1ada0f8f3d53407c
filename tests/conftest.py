import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through the C ABI)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from simpler_amd import engine
    ctx = engine.Context(0)
    yield ctx
    ctx.close()


def gpu_threads():
    """Worker threads for the oracle in GPU tests (box CPU share)."""
    return min(16, os.cpu_count() or 1)
