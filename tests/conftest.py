import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "mc-worldgen-seed-lab_amd"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(HERE, "golden")

# The process can hold two HIP runtimes: torch's bundled one and the system ROCm one that
# libmcbiome_hip.so links.  torch's must be loaded before the library first touches the device, or
# torch later finds no GPU; importing torch here (without initialising the device) fixes the order
# for every test, whichever file runs first.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP engine")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
