import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "era-boojum-cuda_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def gpu_ctx():
    """A live Context on cuda:0 for the whole GPU session (the reference's
    tests create one per test under #[serial]; one process-wide context is
    equivalent here)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but torch.cuda.is_available() is False")
    torch.cuda.set_device(0)
    from boojum_cuda_amd.context import Context
    ctx = Context.create(12, 12)
    yield ctx
    torch.cuda.synchronize()
    ctx.destroy()
