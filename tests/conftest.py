import glob
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "parallel-si_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def load_fixtures():
    out = []
    for f in sorted(glob.glob(os.path.join(GOLDEN, "*.json"))):
        with open(f) as fh:
            out.append(json.load(fh))
    return out


def fixture_game(fx):
    import psi
    g = fx["game"]
    return psi.Game(g["owner"], g["priority"], g["offsets"], g["targets"])


@pytest.fixture(scope="session")
def fixtures():
    return load_fixtures()


@pytest.fixture(scope="session")
def gpu():
    import psi
    psi.init()
    return psi
