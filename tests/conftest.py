import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "quspin.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    """Fail loudly (not skip) when a GPU test cannot reach the native library."""
    try:
        # torch's ROCm runtime must be loaded before libqsgpu initialises HIP,
        # or torch sees no device in the same process
        import torch  # noqa: F401
    except ImportError:
        pass
    import qsgpu
    qsgpu.lib()
    info = qsgpu.device_info()
    print("device:", info, flush=True)
    return info
