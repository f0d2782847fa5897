import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (run with -m gpu)")


@pytest.fixture(scope="session")
def kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    return pyoracle.load()


@pytest.fixture(scope="session")
def lpk():
    from lipika_amd import _lib
    return _lib.load()


@pytest.fixture(scope="session")
def backend():
    from lipika_amd.pipeline import Backend
    be = Backend(0)
    yield be
    be.close()
