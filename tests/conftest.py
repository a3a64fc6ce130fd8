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


@pytest.fixture(params=["valu", "mfma"])
def cmp_engine(request, engine):
    """The engine with one of its two comparison kernels selected."""
    import sdhash_amd
    engine.compare_method(sdhash_amd.CMP_VALU if request.param == "valu" else sdhash_amd.CMP_MFMA)
    yield engine
    engine.compare_method(sdhash_amd.CMP_AUTO)
