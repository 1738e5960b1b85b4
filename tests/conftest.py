import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libtomhip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import tom_oracle
    tom_oracle.lib()
    return tom_oracle


@pytest.fixture(scope="session")
def api():
    from tom_amd import api as _api
    _api.context()
    return _api
