import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test (oracle over many steps)")


@pytest.fixture(scope="session")
def oracle_lib():
    """The CPU oracle (the checker); built on first use if missing."""
    from oracle import pyoracle

    return pyoracle.lib()
