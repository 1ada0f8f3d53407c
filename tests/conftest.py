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


# GPU test files run in dependency order -- forward parity first, then the
# layers built on it -- so that under -x a failure names the lowest layer.
_GPU_ORDER = ["test_gpu_parity.py", "test_gpu_headline.py", "test_gpu_prefilter.py", "test_gpu_host_layer.py", "test_gpu_frontend.py",
              "test_gpu_batch.py", "test_gpu_adjoint.py", "test_gpu_jvp.py", "test_gpu_nee_connect.py"]


def pytest_collection_modifyitems(config, items):
    def rank(item):
        name = os.path.basename(str(item.fspath))
        if name in _GPU_ORDER:
            return _GPU_ORDER.index(name)
        return len(_GPU_ORDER) if name.startswith("test_zz_") else -1  # test_zz_* run last

    items.sort(key=rank)
