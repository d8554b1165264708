import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__
    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def gpu(pkg):
    """The package, with HIP device 0 current; fails (does not skip) without a device."""
    n = pkg.hip.device_count()
    assert n >= 1, "GPU test run without a visible HIP device"
    pkg.hip.check(pkg.hip.lib().hipSetDevice(0), "hipSetDevice")
    return pkg
