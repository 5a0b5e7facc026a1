import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgnd_hip.so on cuda:0)")
    config.addinivalue_line("markers", "slow: long CPU-only statistical checks")


@pytest.fixture(scope="session")
def device():
    from gianduia_amd import Device
    d = Device(0)
    yield d
    d.close()
