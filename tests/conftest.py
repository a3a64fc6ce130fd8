import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libsdbfx.so")


@pytest.fixture(scope="session")
def engine():
    import sdhash_amd
    e = sdhash_amd.Engine(0)
    yield e
    e.close()
