"""GPU parity of the template-search pre-processing (BH_movingAverage_2,
local normalisation, BH_mask3d 'rectangle', the per-chunk body of
BH_templateSearch3d_2.m:532-589 and the chunk driver with its global
variance) against the FP64 oracle. Tolerances: 1e-5 of the maximum for the
moving average, 1e-4 of the maximum for normalised volumes, 1e-4 relative
for the global variance."""
import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.device import DeviceArray
from emclarity_amd.geometry import eulerSearch
from emclarity_amd.masking import BH_mask3d, BH_movingAverage_2
from emclarity_amd.sizing import BH_multi_padVal
from emclarity_amd.template_search import preprocess_chunks, symmetric_pad, template_search
from oracle import emc_oracle as O

pytestmark = pytest.mark.gpu

BP = (1e-3, 600.0, 28.0)
PIXEL = 10.0


def _close(a, b, tol):
    err = float(np.max(np.abs(a - b)))
    assert err <= tol * float(np.max(np.abs(b))), err
    return err


@pytest.mark.parametrize("F", [5, 8, 13])
def test_moving_average_and_local_normalize(ctx, F):
    rng = np.random.default_rng(F)
    a = rng.standard_normal((16, 18, 21)).astype(np.float32)
    _close(BH_movingAverage_2(a, F).get(), O.moving_average(a, F), 1e-5)
    d = DeviceArray.from_numpy(a)
    out = DeviceArray(a.shape, np.float32)
    _lib.call("emc_local_normalize", ctx.handle, d.vp, _lib.i3(d.dims), F, out.vp)
    _close(out.get(), O.local_normalize(a, F), 1e-4)


def test_moving_average_rejects_oversized_feature(ctx):
    d = DeviceArray.from_numpy(np.zeros((8, 8, 8), np.float32))
    with pytest.raises(_lib.EmcError):
        _lib.call("emc_moving_average", ctx.handle, d.vp, _lib.i3(d.dims), 17, d.vp)


@pytest.mark.parametrize("size,radius", [((40, 36, 32), (12, 10, 8)), ((33, 30, 29), (10.5, 8, 9.5))])
def test_rectangle_mask(ctx, size, radius):
    m = BH_mask3d("rectangle", size, radius, [0, 0, 0]).get()
    assert np.max(np.abs(m - O.mask3d_rectangle(size, radius))) < 1e-6


def test_preprocess_chunk(ctx):
    rng = np.random.default_rng(7)
    chunk = (3.0 + 2.0 * rng.standard_normal((28, 32, 36))).astype(np.float32)
    mask = O.mask3d_rectangle((36, 32, 28), (10, 9, 8))
    ref, rsum = O.tsearch_preprocess_chunk(chunk, BP, PIXEL, 8, -1.0, mask)
    d = DeviceArray.from_numpy(chunk)
    dm = DeviceArray.from_numpy(mask.astype(np.float32))
    out = DeviceArray(chunk.shape, np.float32)
    bp = np.array(list(BP) + [PIXEL])
    s = _lib.C.c_double()
    _lib.call("emc_tsearch_preprocess_chunk", ctx.handle, d.vp, _lib.i3(d.dims),
              bp.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)), 0, 8, -1.0, dm.vp, out.vp, _lib.C.byref(s))
    _close(out.get(), ref, 1e-4)
    assert abs(s.value - rsum) <= 1e-5 * float(np.sum(np.abs(ref))), (s.value, rsum)


@pytest.mark.parametrize("k,shape", [(3, (9, 12, 17)), (5, (8, 10, 13)), (7, (6, 9, 11))])
def test_median_filter_matches_oracle(ctx, k, shape):
    """medfilt3 with symmetric padding: the median is one of the inputs, so
    the result is bit-exact; ties and repeated values included."""
    rng = np.random.default_rng(k)
    x = rng.standard_normal(shape).astype(np.float32)
    x[0, :3, :3] = 1.0                       # repeated values at a corner
    x = np.round(x * 8) / 8                  # many ties
    d = DeviceArray.from_numpy(x)
    out = DeviceArray(x.shape, np.float32)
    _lib.call("emc_median_filter3d", ctx.handle, d.vp, _lib.i3(d.dims), k, out.vp)
    assert np.array_equal(out.get(), O.median_filter3(x, k))
    with pytest.raises(_lib.EmcError):
        _lib.call("emc_median_filter3d", ctx.handle, d.vp, _lib.i3(d.dims), 4, out.vp)


def test_preprocess_chunk_median(ctx):
    rng = np.random.default_rng(9)
    chunk = (1.0 + rng.standard_normal((20, 24, 26))).astype(np.float32)
    ref, rsum = O.tsearch_preprocess_chunk(chunk, BP, PIXEL, 6, 1.0, None, med_filt=3)
    d = DeviceArray.from_numpy(chunk)
    out = DeviceArray(chunk.shape, np.float32)
    bp = np.array(list(BP) + [PIXEL])
    s = _lib.C.c_double()
    _lib.call("emc_tsearch_preprocess_chunk", ctx.handle, d.vp, _lib.i3(d.dims),
              bp.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)), 3, 6, 1.0, None, out.vp, _lib.C.byref(s))
    _close(out.get(), ref, 1e-4)


def test_preprocess_driver_and_global_variance(ctx):
    rng = np.random.default_rng(8)
    tomo = rng.standard_normal((30, 36, 40)).astype(np.float32)
    sc, va, it, pre, post, vc = (24, 24, 24), (16, 16, 16), (2, 2, 1), (4, 4, 2), (4, 4, 2), (20, 20, 20)
    rchunks, corners, rgvar, _ = O.tsearch_preprocess(tomo, sc, va, it, pre, post, vc, BP, PIXEL, 8, 1)
    padded = symmetric_pad(tomo, pre, post)
    chunks, gvar = preprocess_chunks(padded, corners, sc, vc, BP, PIXEL, 8, 1)
    assert abs(gvar - rgvar) <= 1e-4 * abs(rgvar), (gvar, rgvar)
    for c, r in zip(chunks, rchunks):
        _close(c.get(), r, 1e-4)


def test_template_search_with_preprocessing(ctx):
    """template_search(preprocess=...) against the oracle's pre-processing
    followed by the oracle chunk search over the same geometry."""
    rng = np.random.default_rng(9)
    t = np.where(rng.uniform(size=(8, 8, 8)) > 0.5, rng.uniform(0.2, 1.0, (8, 8, 8)), 0.0)
    tomo = rng.standard_normal((24, 24, 40)).astype(np.float32)
    grid = eulerSearch("C1", 30, 30, 30, 30)
    size_chunk, valid, n_iters = (32, 32, 32), (20, 24, 24), (2, 1, 1)
    pre, post, vc = (6, 4, 4), (6, 4, 4), (28, 28, 28)
    pp = dict(valid_calc=vc, bp_vals=BP, pixel=PIXEL, stats_radius=8, should_be_ctf=1)
    peak, ang = template_search(tomo, t.astype(np.float32), grid, size_chunk, valid, n_iters, pre, post,
                               preprocess=pp)
    rchunks, corners, _, _ = O.tsearch_preprocess(tomo, size_chunk, valid, n_iters, pre, post, vc, BP, PIXEL, 8, 1)
    padded_shape = (24 + 8, 24 + 8, 40 + 12)
    pad = BH_multi_padVal((8, 8, 8), size_chunk)
    vA = BH_multi_padVal(valid, size_chunk)
    trim = BH_multi_padVal(size_chunk, valid)
    rp = np.zeros(padded_shape)
    ra = np.zeros(padded_shape)
    thetas = sorted(set(grid[:, 1]), key=list(grid[:, 1]).index)
    first = True
    for th in thetas:
        idx = np.flatnonzero(grid[:, 1] == th)
        for k, ((cx, cy, cz), sub) in enumerate(zip(corners, rchunks)):
            z0, y0, x0 = cz + vA[0][2], cy + vA[0][1], cx + vA[0][0]
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
    assert np.max(np.abs(peak - rp)) < 2e-4 * np.max(np.abs(rp))
    assert np.mean(ang == ra) > 0.999
