import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_emdee.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle library (test infrastructure), built on demand."""
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    from emdee_amd.emdee import EmDeeLibrary
    return EmDeeLibrary(ORACLE_SO)


@pytest.fixture(scope="session")
def gpu():
    """The HIP product library; fails (does not skip) when it is missing."""
    os.environ.setdefault("EMDEE_QUIET", "1")
    import emdee_amd
    return emdee_amd.load()
