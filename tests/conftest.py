import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def engine():
    """The HIP engine; GPU tests fail loudly if it is not built or no GPU."""
    from kmos_amd import _lib
    _lib.load()
    assert _lib.device_count() > 0, "no GPU visible"
    from kmos_amd import engine as eng
    return eng
