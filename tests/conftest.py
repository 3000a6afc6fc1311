import os
import sys
from pathlib import Path

import pytest

# a kernel that misses its exit condition is stopped by the library's device
# watchdog after this many seconds instead of the default
os.environ.setdefault("NOMAD_WATCHDOG_SECONDS", "20")

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")


def _gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _gpu_available():
        pytest.skip("no GPU visible")
    from nomad_amd import lib

    lib.hip()  # fails loudly if the HIP library is not built
    return lib


@pytest.fixture(scope="session")
def small():
    """Config-1 shaped data (2,000 x 1,000, 100,000 cells) at full size."""
    from nomad_amd import lib

    tr, te = lib.synth(lib.preset(1, 1.0))
    return tr, te


@pytest.fixture(scope="session")
def tiny():
    """A 300 x 120 uniform problem for exhaustive bitwise checks."""
    from nomad_amd import lib

    tr, te = lib.synth(lib.preset(1, 0.15))
    return tr, te
