import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libnphys_hip.so")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle

    # make is a no-op when the library is newer than its sources and the shared header; where
    # there is no compiler (a machine that only runs tests) the prebuilt library is used
    try:
        oracle.build()
    except (OSError, subprocess.CalledProcessError):
        if not os.path.exists(oracle.LIB_PATH):
            raise
    return oracle


@pytest.fixture(scope="session")
def nphys():
    from nphysics_amd import world

    world.load_library()
    return world
