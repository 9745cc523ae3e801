import os
import sys

import pytest

# Load torch (and the HIP runtime it ships) before libt3ns_amd.so so that
# one HIP runtime serves both in this process.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

NETW = os.path.join(ROOT, "tests", "golden", "networks")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def netw_text(name):
    """Network fixture: a bare name is a makenetwork.py output under
    tests/golden/networks, a path is relative to tests/golden."""
    path = os.path.join(GOLDEN, name) if "/" in name else os.path.join(NETW, name)
    with open(path + ".netw") as f:
        return f.read()


def have_gpu():
    try:
        from t3ns_amd import lib
        return lib().t3ns_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not have_gpu():
        pytest.fail("GPU test run without a visible HIP device or without libt3ns_amd.so")
    return True
