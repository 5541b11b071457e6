import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT), str(ROOT / "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import pgwload  # noqa: E402

pgwload.load()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libpgw_hip.so")


@pytest.fixture(scope="session")
def synth():
    from polgraw_allsky_amd import synth as s
    return s


@pytest.fixture(scope="session")
def abi():
    from polgraw_allsky_amd import abi as a
    return a


@pytest.fixture(scope="session")
def orc():
    import orc as o
    return o
