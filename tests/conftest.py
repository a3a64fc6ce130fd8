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


@pytest.fixture(params=["valu", "mfma", "mfma_i8"])
def cmp_engine(request, engine):
    """The engine with one of its comparison kernels selected: VALU
    popcount, MFMA on FP4 bit elements, MFMA on int8 byte planes."""
    import sdhash_amd
    engine.compare_method({"valu": sdhash_amd.CMP_VALU, "mfma": sdhash_amd.CMP_MFMA,
                           "mfma_i8": sdhash_amd.CMP_MFMA_I8}[request.param])
    yield engine
    engine.compare_method(sdhash_amd.CMP_AUTO)
