"""Shared fixtures: package / oracle import helpers and the gpu marker."""
import importlib
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))
ORACLE = ROOT / "oracle"
if str(ORACLE) not in sys.path:
    sys.path.insert(0, str(ORACLE))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs a HIP device (MI355X) and the built library")


def pkg():
    return importlib.import_module("kern-interp-dev_amd")


def geometry():
    return importlib.import_module("kern-interp-dev_amd.geometry")


@pytest.fixture(scope="session")
def ki():
    return pkg()


@pytest.fixture(scope="session")
def geo():
    return geometry()


@pytest.fixture(scope="session")
def oracle():
    import ki_oracle
    return ki_oracle
