import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests", "emu"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_oracle():
    lib = os.path.join(ROOT, "oracle", "libmtgp_oracle.so")
    src = os.path.join(ROOT, "oracle", "mtgp_oracle.c")
    if not os.path.exists(lib) or os.path.getmtime(lib) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])


@pytest.fixture(scope="session")
def oracle():
    _ensure_oracle()
    from pyoracle import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def mtgp():
    import mtgp_amd
    return mtgp_amd
