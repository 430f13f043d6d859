import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def golden_problems():
    with open(os.path.join(GOLDEN, "problems.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_fixtures():
    with open(os.path.join(GOLDEN, "fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def golden_rng():
    with open(os.path.join(GOLDEN, "rng.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def replay_stats():
    with open(os.path.join(GOLDEN, "replay_stats.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle_fma():
    from oracle_lib import Oracle
    return Oracle(variant="fma")


@pytest.fixture(scope="session")
def oracle_ld():
    from oracle_lib import Oracle
    return Oracle(variant="ld")


@pytest.fixture(scope="session")
def engine():
    import feyntrop_amd
    e = feyntrop_amd.Engine(0)
    yield e
    e.close()
