"""The MEX gateway sources (emclarity_amd/mex) compile as C11 against
emc_hip.h and the documented MATLAB C Matrix API (declarations in
tests/mex_api/mex.h), with every warning an error: argument types of each
C-ABI call they make are checked by the compiler. MATLAB is not here, so
nothing is linked or run."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GATEWAYS = sorted(glob.glob(os.path.join(ROOT, "emclarity_amd", "mex", "*.c")))


def test_every_gateway_is_present():
    names = {os.path.basename(g) for g in GATEWAYS}
    assert names == {"mexXform3d_amd.c", "mexSF3D_amd.c", "mexCTF_amd.c", "mexFFT_amd.c",
                     "BH_multi_xcf_Translational.c"}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
@pytest.mark.parametrize("src", GATEWAYS, ids=os.path.basename)
def test_gateway_compiles(src):
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "tests", "mex_api"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
