import ctypes
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_BIN = os.path.join(ORACLE_DIR, "build", "sa_oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "build", "liborc.so")

if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_oracle():
    if not (os.path.exists(ORACLE_BIN) and os.path.exists(ORACLE_LIB)):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


@pytest.fixture(scope="session")
def oracle_bin():
    _ensure_oracle()
    return ORACLE_BIN


@pytest.fixture(scope="session")
def orc():
    _ensure_oracle()
    return ctypes.CDLL(ORACLE_LIB)


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
