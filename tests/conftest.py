import json
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "metal-reax_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(REPO, "tests", "golden", "fixtures.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def ffield_text():
    with open(os.path.join(REPO, "tests", "golden", "ffield.cho_standin")) as fh:
        return fh.read()
