import os
import sys

# The oracle's OpenMP loops run on grids of a few thousand cells in these
# tests: one thread each, so that parallel test workers do not oversubscribe
# the cores (an explicit OMP_NUM_THREADS is kept).
os.environ.setdefault("OMP_NUM_THREADS", "1")

import pytest  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
