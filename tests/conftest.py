import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "exposure-render_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: long-running statistical check")


@pytest.fixture(scope="session")
def lib():
    from exposure_render import capi
    return capi.load()


@pytest.fixture(scope="session")
def orc_lib():
    import oracle
    return oracle.load()


@pytest.fixture()
def fresh(lib):
    """A clean registry (IDs count from 0) before and after a GPU test."""
    import exposure_render as er
    er.shutdown()
    yield
    er.shutdown()
