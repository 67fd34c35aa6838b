"""torch.autograd.gradcheck of the adjoint behind AlignScore, in a child
process (the child loads the package's HIP library before importing torch;
torch's own GPU runtime is never initialised)."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_align_score_gradcheck():
    r = subprocess.run([sys.executable, "-m", "emclarity_amd.autograd"], cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "gradcheck passed" in r.stdout
