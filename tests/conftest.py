import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_SO = os.path.join(ROOT, "oracle", "libsrx_cpu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def cpu_lib():
    """The CPU oracle (test infrastructure), built from oracle/ if missing."""
    from salviarenderer_amd import api

    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return api.Lib(ORACLE_SO)


@pytest.fixture(scope="session")
def hip_lib():
    """The product library; fails (does not skip) when it is not built."""
    from salviarenderer_amd import hip_lib as load

    return load()
