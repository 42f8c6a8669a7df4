import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
for p in (REPO / "tests", REPO / "oracle", REPO / "heterogeneous-clue_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not oracle.LIB.exists():
        oracle.build(with_reference=False)
    return oracle.Oracle()


@pytest.fixture(scope="session")
def dev():
    import clue_hip
    d = clue_hip.Device(0)
    yield d
    d.close()
