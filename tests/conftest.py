import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU runs")


def gpu_available():
    try:
        import mps_amd
        lib = mps_amd.load(3)
        return lib.mps_hip_init(0) == 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def lib3():
    import mps_amd
    lib = mps_amd.load(3)
    rc = lib.mps_hip_init(0)
    if rc != 0:
        raise RuntimeError("GPU library did not initialise: %s" % lib.mps_hip_error_string(rc).decode())
    lib.set_option("exit_on_error", 0)
    return lib


@pytest.fixture(scope="session")
def lib3s():
    """The 3-D library in the serial program's layout (with friction)."""
    import mps_amd
    lib = mps_amd.load(3, serial=True)
    rc = lib.mps_hip_init(0)
    if rc != 0:
        raise RuntimeError("GPU library did not initialise: %s" % lib.mps_hip_error_string(rc).decode())
    lib.set_option("exit_on_error", 0)
    return lib


@pytest.fixture(scope="session")
def lib2():
    import mps_amd
    lib = mps_amd.load(2)
    rc = lib.mps_hip_init(0)
    if rc != 0:
        raise RuntimeError("GPU library did not initialise: %s" % lib.mps_hip_error_string(rc).decode())
    lib.set_option("exit_on_error", 0)
    return lib
