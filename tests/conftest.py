import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def ref_cases():
    return json.loads((GOLDEN / "reference_cases.json").read_text())["suites"]


@pytest.fixture(scope="session")
def golden_tables():
    return json.loads((GOLDEN / "golden_tables.json").read_text())


@pytest.fixture(scope="session")
def golden_trials():
    return json.loads((GOLDEN / "golden_trials.json").read_text())


def hex_to_words(h):
    return np.array([int(h[16 * i:16 * i + 16], 16) for i in range(len(h) // 16)], dtype=np.uint64)


def cdiv(a, b):
    """C/C++ integer division (truncation toward zero)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q
