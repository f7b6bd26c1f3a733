import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """Builds the product library and the oracle when they are missing."""
    from kittlespt_amd import abi, build
    import oracle
    if not os.path.exists(abi.LIB_PATH):
        build.build()
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    yield
