import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "deepmd-kit_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: long CPU test")


def golden_path(name):
    return os.path.join(GOLDEN, name)


@pytest.fixture(scope="session")
def cc_goldens():
    with open(golden_path("cc_goldens.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def cc_geometry():
    with open(golden_path("cc_geometry.json")) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def sea_yaml():
    import yaml
    with open(golden_path("deeppot_sea.yaml")) as fh:
        return yaml.safe_load(fh)


@pytest.fixture(scope="session")
def sea_testcase():
    import yaml
    with open(golden_path("deeppot-testcase.yaml")) as fh:
        return yaml.safe_load(fh)
