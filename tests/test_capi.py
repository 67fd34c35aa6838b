"""C ABI checks that need no GPU: the library loads, exports every symbol
include/emc_hip.h declares, and the host-only entry points (FP64 geometry,
sizing) agree with the oracle."""
import ctypes as C
import re

import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.geometry import BH_defineMatrix, eulerSearch
from emclarity_amd import sizing
from oracle import emc_oracle as O


def declared_symbols():
    text = _lib.HEADER_PATH.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(emc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes table covers the same set
    assert sorted(_lib.SIGNATURES) == names


def test_version_and_error_path():
    lib = _lib.load()
    assert b"gfx950" in lib.emc_version()
    # invalid arguments are reported, never exit the process
    st = lib.emc_define_matrix(_lib.d_arr([0, 0, 0]), 9, 0, (C.c_double * 9)())
    assert st == 1 and b"convention" in lib.emc_last_error()


@pytest.mark.parametrize("conv", ["Bah", "TILT", "SPIDER", "Helical", "IMOD"])
@pytest.mark.parametrize("direction", ["forward", "inv", "forwardVector", "invVector"])
def test_define_matrix_matches_oracle(conv, direction, rng):
    for _ in range(4):
        a = rng.uniform(-180, 180, 3)
        if conv == "TILT":
            a = a[:1]
        assert np.allclose(BH_defineMatrix(a, conv, direction), O.define_matrix(a, conv, direction),
                           atol=1e-15)


@pytest.mark.parametrize("args", [("C1", 180, 10, 180, 10), ("C1", 15, 3, 15, 3), ("C4", 90, 15, 180, 20),
                                  ("D2", 180, 30, 180, 30), ("O", 90, 10, 90, 10), ("C1", -30, 10, 10, 5)])
def test_euler_search_matches_oracle(args):
    a = eulerSearch(*args)
    b = O.euler_search(*args).orientations()
    assert a.shape == b.shape and np.allclose(a, b, atol=1e-12)


def test_sizing_helpers():
    # padVal: odd/even origin rule (BH_multi_padVal.m)
    assert np.array_equal(sizing.BH_multi_padVal([63] * 3, [64] * 3)[:2], O.pad_val([63] * 3, [64] * 3)[:2])
    assert np.array_equal(sizing.BH_multi_padVal([48] * 3, [480] * 3), O.pad_val([48] * 3, [480] * 3))
    # config-1 chunking, worked by hand from BH_multi_iterator 'convolution':
    # border 30 + 2*36 = 102; x: 480 scores 480/143 (best), nIters floor(512/276)+1 = 2;
    # z: 480 -> 480/122, nIters floor(256/276)+1 = 1
    r = sizing.iterator_convolution([512] * 3, [512, 512, 256], [48] * 3, [36] * 3)
    assert np.array_equal(r["chunkSize"], [480, 480, 480])
    assert np.array_equal(r["nIters"], [2, 2, 1])
    assert np.array_equal(r["tomoPre"], [102] * 3) and np.array_equal(r["tomoPost"], [143, 143, 122])
    w, c, m, pw, pc = sizing.BH_multi_validArea([63] * 3, [14] * 3, 1.0)
    assert np.array_equal(c, [64] * 3) and np.array_equal(pc[0], [1, 1, 1])
    ind, pad, sh = sizing.BH_isWindowValid([200, 200, 100], [64] * 3, [20] * 3, [100.3, 50.0, 40.75])
    assert np.allclose(sh, [0.3, 0.0, 0.75]) and np.array_equal(ind[:, 0], [68, 131])
    assert sizing.BH_isWindowValid([200, 200, 100], [64] * 3, [20] * 3, [3.0, 50.0, 40.0])[0] == "noUse"


def test_symmetry_operators_match_the_oracle():
    """Host point groups (closure of generators) against the oracle's
    (enumeration of vertex frames), as sets; D<n> is the n C_n operators."""
    from emclarity_amd.geometry import symmetry_matrices
    for name in ("C5", "D3", "O", "I", "I2"):
        a, b = symmetry_matrices(name), O.symmetry_matrices(name)
        assert len(a) == len(b)
        assert all(any(np.max(np.abs(x - y)) < 1e-9 for y in b) for x in a), name
    with pytest.raises(ValueError):
        symmetry_matrices("O", "Helical")
