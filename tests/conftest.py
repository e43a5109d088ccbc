import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def emu_lib():
    path = os.path.join(ROOT, "tests", "emu", "libpb_emu.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.dirname(path)], check=True, stdout=subprocess.DEVNULL)
    import ctypes
    return ctypes.CDLL(path)
