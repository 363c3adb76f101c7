import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X, gfx950)")


@pytest.fixture(scope="session")
def rt():
    """The product package (its directory name starts with a digit)."""
    return importlib.import_module("2dradiation_amd")


@pytest.fixture(scope="session")
def O():
    """The CPU oracle (checker only)."""
    from oracle import oracle
    return oracle
