import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: long CPU runs")


@pytest.fixture(scope="session", autouse=True)
def _built_oracle():
    """The plain-C oracle is a few hundred lines of C: build it if needed."""
    from oracle import orcpy
    orcpy.build()
    yield
