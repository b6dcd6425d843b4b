import os
import sys

import pytest

# PyTorch-ROCm ships its own HIP runtime with the same soname as /opt/rocm's.
# Whichever loads first serves the whole process, and torch fails to find the
# GPU when the library under test loaded /opt/rocm's first, so torch goes first.
import torch  # noqa: F401,E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    return np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))
