import glob
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def gpu_present():
    return os.path.exists("/dev/kfd") and bool(glob.glob("/dev/dri/renderD*"))


@pytest.fixture(scope="session")
def gpu():
    if not gpu_present():
        pytest.skip("no GPU visible")
    from ansatzevolve_amd import load_library
    return load_library()
