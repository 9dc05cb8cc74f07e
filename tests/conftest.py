import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (built from oracle/ if its shared library is missing)."""
    from oracle import oracle as O
    if not os.path.exists(O.LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    O.lib()
    return O


@pytest.fixture(scope="session")
def exa():
    """The HIP product library; fails (never skips) when it cannot be loaded."""
    from exahype_amd import capi
    capi.lib()
    return capi
