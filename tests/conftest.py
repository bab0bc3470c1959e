import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def gpu():
    import qz_amd
    if not qz_amd.lib().qzbsc_gpu_available():
        pytest.fail("HIP device not available to libqzbsc.so")
    return qz_amd
