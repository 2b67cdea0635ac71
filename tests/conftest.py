import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
for p in (ROOT / "bachelor-thesis_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libpmarch.so on device 0)")


@pytest.fixture(scope="session")
def marcher():
    import pmarch
    m = pmarch.Marcher(0)
    yield m
    m.close()
