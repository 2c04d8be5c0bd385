import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through liborqg.so)")


@pytest.fixture(scope="session")
def keys():
    from orq_amd import keys_from_seed

    return keys_from_seed(0x0A11CE01)
