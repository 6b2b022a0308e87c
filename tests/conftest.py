import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libophip on cuda:0)")


@pytest.fixture(scope="session")
def hip():
    """The ophip binding with the device context initialised."""
    from oplib_amd import ophip
    ophip.init()
    return ophip
