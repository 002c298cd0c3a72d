import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG = ROOT / "boinc-einstein-bpr4-v3d_amd"
for path in (str(PKG), str(ROOT / "tests")):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs through the C ABI of libbrp_hip.so")
    config.addinivalue_line("markers", "slow: full-size (N = 2^22) cases")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib

    oracle_lib.load()
    return oracle_lib


@pytest.fixture(scope="session")
def brp():
    import brp_hip

    brp_hip.load()
    return brp_hip
