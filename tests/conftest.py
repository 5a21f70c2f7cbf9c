import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) ROCm device")


@pytest.fixture(scope="session")
def gps_lib():
    """libgps.so loaded through ctypes; built first if it is missing."""
    from gpuprefixsums_amd import _lib
    if not _lib.LIB_PATH.exists():
        from gpuprefixsums_amd import build
        build.build_lib()
    return _lib.load()
