import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    # torch's HIP runtime (it provides device buffers to some GPU tests)
    # must initialise before libxbg's: once libxbg has opened the device,
    # torch reports no GPU
    try:
        import torch

        torch.cuda.is_available()
    except Exception:
        pass


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401
        return bool(torch.cuda.is_available())
    except Exception:
        return False


@pytest.fixture(scope="session")
def need_gpu():
    if not gpu_available():
        pytest.skip("no GPU visible")
