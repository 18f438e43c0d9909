import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libsg_hip.so kernels)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle module (built on first use)."""
    from oracle import pysgcpu
    pysgcpu.load()
    return pysgcpu


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_known_answers.json")) as f:
        return json.load(f)
