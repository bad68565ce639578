import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_LIB = os.path.join(ROOT, "oracle", "libzmpm_oracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from ziran2020_amd._abi import ZmpmLib
    if not os.path.exists(ORACLE_LIB):
        subprocess_build_oracle()
    return ZmpmLib(ORACLE_LIB)


def subprocess_build_oracle():
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


@pytest.fixture(scope="session")
def gpu():
    """The HIP library; a GPU test fails (never falls back) if it is missing."""
    import ziran2020_amd
    return ziran2020_amd.load_library()
