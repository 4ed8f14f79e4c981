import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libsloth_hip.so")


@pytest.fixture(scope="session")
def orc():
    """The FP64 CPU oracle (test infrastructure)."""
    from oracle import oracle as O

    O.load()
    return O
