import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: full-size workload")


@pytest.fixture(scope="session")
def gpu():
    """Returns the device description; fails (does not skip) when the HIP library is not built."""
    import graphone_amd as g1

    info = g1.device_info()
    assert info["devices"] >= 1
    return info
