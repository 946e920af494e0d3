import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "vapoursynth-descalemvc_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "reference: needs oracle/_ref (reference engine build)")


@pytest.fixture(scope="session")
def oracle():
    from support import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from support import Reference, reference_available
    if not reference_available():
        pytest.skip("oracle/_ref/libdsmvc_ref.so not built (reference sources absent)")
    return Reference()
