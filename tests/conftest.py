import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP path on an MI355X")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(GOLDEN, "reference_tests.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def ctx():
    from lehrfempp_amd import capi
    c = capi.Context(0)
    yield c
    c.close()
