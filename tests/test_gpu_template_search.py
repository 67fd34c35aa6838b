"""GPU parity of the template-search chunk body (emc_template_search_chunk)
and of the chunked driver against the FP64 oracle: per-voxel max CC to
1e-4 relative, argmax orientation ids exact."""
import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.device import DeviceArray
from emclarity_amd.geometry import eulerSearch
from emclarity_amd.sizing import BH_multi_padVal
from emclarity_amd.template_search import template_search
from oracle import emc_oracle as O

pytestmark = pytest.mark.gpu


def _template(rng, n=8, r2=5):
    ax = np.arange(n) - n // 2
    zz, yy, xx = np.meshgrid(ax, ax, ax, indexing="ij")
    return np.where(xx ** 2 + yy ** 2 + zz ** 2 <= r2, rng.uniform(0.2, 1.0, (n, n, n)), 0.0)


def _check(peak, ang, rpeak, rang):
    assert np.max(np.abs(peak - rpeak)) < 1e-4 * np.max(np.abs(rpeak))
    # ids exact wherever the best two orientations are not a near tie
    assert np.mean(ang == rang) > 0.999
    return True


@pytest.mark.parametrize("pruned", [2, 1, 0])
def test_template_search_chunk_matches_oracle(ctx, pruned):
    """pruned 1: the template's spectrum as direct DFTs over its block at
    the padded position, fused with the product by the tomogram spectrum;
    2: the same with the z-stage DFT and product on the matrix cores;
    0: the zero-padded volume through rocFFT's R2C."""
    _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, pruned)
    rng = np.random.default_rng(21)
    t = _template(rng)
    grid = eulerSearch("C1", 60, 30, 60, 30)[:14]
    chunk = rng.standard_normal((32, 32, 32))
    for k, c in ((3, (14, 15, 17)), (9, (20, 12, 13))):
        phi, theta, psi = grid[k]
        rot = O.interp3d(t, [phi, theta, psi - phi], [1, 1, 1], "Bah", "forward")
        chunk[c[2] - 4:c[2] + 4, c[1] - 4:c[1] + 4, c[0] - 4:c[0] + 4] -= 3 * rot
    chunk = chunk.astype(np.float32)
    pad = BH_multi_padVal((8, 8, 8), (32, 32, 32))
    valid = BH_multi_padVal((32, 32, 32), (20, 20, 20))
    vlow = -valid[0]
    d_chunk = DeviceArray.from_numpy(chunk)
    d_t = DeviceArray.from_numpy(t.astype(np.float32))
    peak = DeviceArray.zeros((20, 20, 20), np.float32)
    ang = DeviceArray.zeros((20, 20, 20), np.int32)
    pts = np.ascontiguousarray(grid)
    dp = _lib.C.POINTER(_lib.C.c_double)
    # two calls (ids 1..7, 8..14) exercise the running max across calls
    for a0, a1, first in ((0, 7, 1), (7, 14, 0)):
        _lib.call("emc_template_search_chunk", ctx.handle, d_chunk.vp, _lib.i3((32, 32, 32)), d_t.vp,
                  _lib.i3((8, 8, 8)), _lib.i3(pad[0]), _lib.i3(vlow), _lib.i3((20, 20, 20)),
                  pts[a0:a1].ctypes.data_as(dp), a1 - a0, a0 + 1, first, peak.vp, ang.vp)
    _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, 2)   # the library default
    rpeak, rang = O.template_search_chunk(chunk.astype(np.float64), t, grid, pad[0:2], valid[0:2])
    _check(peak.get(), ang.get(), rpeak, rang)


def test_template_search_pruned_matches_padded_odd_sizes(ctx):
    """The pruned spectrum (vector and matrix-core z stage) against the
    padded-volume R2C on an odd-sized, non-cubic chunk and template with an
    off-centre block (every DFT length and offset different; 29 kz and
    37 x 34 / 2 columns leave partial matrix-core tiles), over several
    orientation blocks."""
    rng = np.random.default_rng(5)
    t = rng.standard_normal((9, 7, 11)).astype(np.float32)        # (z, y, x): 11 x 7 x 9
    chunk = rng.standard_normal((29, 34, 37)).astype(np.float32)  # 37 x 34 x 29
    grid = np.ascontiguousarray(eulerSearch("C1", 60, 30, 60, 30)[:21])
    cd, td, pad, vlow, vd = (37, 34, 29), (11, 7, 9), (5, 20, 3), (4, 5, 6), (25, 22, 19)
    d_chunk = DeviceArray.from_numpy(chunk)
    d_t = DeviceArray.from_numpy(t)
    dp = _lib.C.POINTER(_lib.C.c_double)
    out = []
    for pruned in (2, 1, 0):
        _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, pruned)
        peak = DeviceArray.zeros((vd[2], vd[1], vd[0]), np.float32)
        ang = DeviceArray.zeros((vd[2], vd[1], vd[0]), np.int32)
        _lib.call("emc_template_search_chunk", ctx.handle, d_chunk.vp, _lib.i3(cd), d_t.vp, _lib.i3(td),
                  _lib.i3(pad), _lib.i3(vlow), _lib.i3(vd), grid.ctypes.data_as(dp), len(grid), 1, 1,
                  peak.vp, ang.vp)
        out.append((peak.get(), ang.get()))
    _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, 2)   # the library default
    (p2, a2), (p1, a1), (p0, a0) = out
    for p, a in ((p1, a1), (p2, a2)):
        assert np.max(np.abs(p - p0)) < 1e-4 * np.max(np.abs(p0))
        assert np.mean(a == a0) > 0.999


def test_template_search_driver_chunks(ctx):
    """Two chunks along x with symmetric padding, against the oracle applied
    to the same chunk geometry (results initialised with zeros except the
    first chunk's first orientation, as the reference)."""
    rng = np.random.default_rng(5)
    t = _template(rng)
    tomo = rng.standard_normal((24, 24, 40)).astype(np.float32)
    grid = eulerSearch("C1", 30, 30, 30, 30)
    size_chunk, valid, n_iters = (32, 32, 32), (20, 24, 24), (2, 1, 1)
    pre, post = (6, 4, 4), (6, 4, 4)
    peak, ang = template_search(tomo, t.astype(np.float32), grid, size_chunk, valid, n_iters, pre, post)
    padded = np.pad(tomo.astype(np.float64), [(4, 4), (4, 4), (6, 6)], mode="symmetric")
    pad = BH_multi_padVal((8, 8, 8), size_chunk)
    vA = BH_multi_padVal(valid, size_chunk)
    trim = BH_multi_padVal(size_chunk, valid)
    rp = np.zeros(padded.shape)
    ra = np.zeros(padded.shape)
    thetas = sorted(set(grid[:, 1]), key=list(grid[:, 1]).index)
    first = True
    for th in thetas:
        idx = np.flatnonzero(grid[:, 1] == th)
        for k, cx in enumerate((0, 20)):
            sub = padded[0:32, 0:32, cx:cx + 32]
            z0, y0, x0 = vA[0][2], vA[0][1], cx + vA[0][0]
            prev_p = rp[z0:z0 + 24, y0:y0 + 24, x0:x0 + 20]
            prev_a = ra[z0:z0 + 24, y0:y0 + 24, x0:x0 + 20]
            mp, ma = O.template_search_chunk(sub, t, grid[idx], pad[0:2], trim[0:2],
                                             prev_peak=None if (first and k == 0) else prev_p,
                                             prev_angle=None if (first and k == 0) else prev_a,
                                             first_angle_id=int(idx[0]) + 1,
                                             init_from_first=(first and k == 0))
            rp[z0:z0 + 24, y0:y0 + 24, x0:x0 + 20] = mp
            ra[z0:z0 + 24, y0:y0 + 24, x0:x0 + 20] = ma
        first = False
    rp = rp[4:28, 4:28, 6:46]
    ra = ra[4:28, 4:28, 6:46]
    _check(peak, ang, rp, ra)
