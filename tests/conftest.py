"""Test configuration: the `gpu` marker and shared fixtures loaders."""

import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


def golden_json(name):
    return json.loads((GOLDEN / name).read_text())


def golden_npz(name):
    return np.load(GOLDEN / name, allow_pickle=False)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.library()
    return O
