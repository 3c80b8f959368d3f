import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "endless-games-of-war_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(ROOT, "tests", "golden", "goldens.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def eng():
    """The engine on device 0. A missing libwar_engine.so is an error (OSError), never a skip;
    only the absence of a GPU (status 50 from war_init) skips."""
    from war_amd import Engine, WarError
    try:
        e = Engine(0)
    except WarError as ex:
        if ex.status == 50:
            pytest.skip("no GPU: %s" % ex)
        raise
    yield e
