import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvthip.so on cuda:0)")


@pytest.fixture(scope="session")
def engine():
    """One engine for the whole GPU session (device 0)."""
    from vitrum_amd import native
    eng = native.Engine(0)
    yield eng
    eng.close()
