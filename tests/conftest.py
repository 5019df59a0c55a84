import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libwfhip.so")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_ensure_oracle()


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def wflib():
    """The HIP library; GPU tests fail loudly if it is missing."""
    from weldformgpu_amd import _abi
    return _abi.load()
