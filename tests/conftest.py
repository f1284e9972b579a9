import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    import tws_oracle
    tws_oracle.lib()
    return tws_oracle


@pytest.fixture(scope="session")
def tws():
    import terrainwatersim_amd
    terrainwatersim_amd.lib()
    return terrainwatersim_amd


@pytest.fixture(scope="session")
def gpu(tws):
    if tws.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return tws
