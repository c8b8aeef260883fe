import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def skl():
    """The initialised library on GPU 0.  Fails (never skips) when the HIP
    library or the GPU is missing: GPU tests must run the native path."""
    import skelcl_amd as skelcl
    skelcl.init()
    print("device:", skelcl.device_info()["name"], skelcl.device_info()["arch"], flush=True)
    return skelcl
