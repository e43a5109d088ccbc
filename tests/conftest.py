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


def _make(directory, target):
    """Run make when a compiler is present (it rebuilds a library whose
    sources changed: the Makefiles list them); on a box without one the
    prebuilt library is used as it is."""
    import shutil
    if shutil.which("make") and shutil.which(os.environ.get("CXX", "g++")):
        subprocess.run(["make", "-C", directory, target], check=True, stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle as O
    _make(os.path.join(ROOT, "oracle"), "liboracle.so")
    O.build()
    return O


@pytest.fixture(scope="session")
def emu_lib():
    path = os.path.join(ROOT, "tests", "emu", "libpb_emu.so")
    _make(os.path.dirname(path), "libpb_emu.so")
    import ctypes
    return ctypes.CDLL(path)
