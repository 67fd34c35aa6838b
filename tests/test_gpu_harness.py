"""The plain C harness (no Python, no MATLAB) runs the C ABI end to end."""
import subprocess

import pytest

from emclarity_amd.csrc.build import HARNESS_BIN

pytestmark = pytest.mark.gpu


def test_c_harness_runs():
    assert HARNESS_BIN.exists(), "build the library first (python -m emclarity_amd.csrc.build)"
    r = subprocess.run([str(HARNESS_BIN)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "harness OK" in r.stdout
