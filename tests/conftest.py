import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-presolver_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def _ensure_built(path, directory):
    if not os.path.exists(path):
        subprocess.run(["make", "-C", directory, "-j", "8"], check=True,
                       stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    _ensure_built(os.path.join(ORACLE, "build", "liboracle.so"), ORACLE)
