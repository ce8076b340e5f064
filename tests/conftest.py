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
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU run")


def pytest_collection_modifyitems(config, items):
    # a GPU test that cannot reach a GPU must fail loudly, never fall back
    pass


@pytest.fixture(scope="session")
def ora():
    import oracle_bind
    return oracle_bind.load()
