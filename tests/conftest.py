import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def built():
    """Loads the in-tree library, building it only when it is missing."""
    from thor_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        from thor_amd import build
        build.build()
    capi.lib()
    from oracle import oracle
    oracle.build()
    return capi
