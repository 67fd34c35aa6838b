import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests proper")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session (fails loudly without the
    library or a device: GPU tests never fall back to the CPU)."""
    from emclarity_amd.device import default_context
    return default_context()


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_tlt(name="riboNew05_ali1_ctf.tlt"):
    return np.loadtxt(os.path.join(GOLDEN, "tlt", name))
