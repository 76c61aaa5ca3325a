import os
import sys

import pytest

# The oracle is an OpenMP library with a barrier between passes; on the small
# grids of most tests several threads per process (more so with pytest-xdist
# workers) spend their time spinning at those barriers.  One thread unless
# the environment asks for more (the GPU host sets its own CPU share).
os.environ.setdefault("OMP_NUM_THREADS", "1")

# torch before libmpdx.so: torch carries its own HIP runtime, and a process
# that first loads the system one (through libmpdx.so, which links it) cannot
# initialise torch's GPU context afterwards ("No HIP GPUs are available");
# loaded in this order libmpdx.so binds to the copy torch brought.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
