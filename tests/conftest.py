import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name):
    path = os.path.join(GOLDEN, name + ".npz")
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name}.npz not generated")
    return np.load(path)


@pytest.fixture(scope="session")
def pg():
    """The product bindings; builds the native library if it is not there."""
    import pine_amd
    if not os.path.exists(pine_amd.LIB_PATH):
        from pine_amd import build
        build.build()
    pine_amd.lib()
    return pine_amd


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def cuda(pg):
    """Skips unless a HIP device is visible to the library's runtime."""
    from pine_amd import hip
    if hip.device_count() < 1:
        pytest.skip("no GPU")
    return hip
