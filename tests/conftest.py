import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def lglib():
    from light_garden_amd import abi
    if not os.path.exists(abi.LIB_PATH):
        from light_garden_amd import build
        build.build()
    return abi.load()


@pytest.fixture(scope="session")
def ctx(lglib):
    import light_garden_amd as lg
    c = lg.Context(0)
    yield c
    c.close()
