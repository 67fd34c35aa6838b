"""GPU parity of the template-search peak extraction (BH_mask3d 'sphere',
the greedy erase loop of BH_templateSearch3d_2.m:1100-1275 and the driver
around it) against the FP64 oracle: mask 1e-5 absolute, peak positions
1e-4 voxel, angles exact, scores 1e-5 relative."""
import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.device import DeviceArray
from emclarity_amd.masking import BH_mask3d
from emclarity_amd.template_search import extract_peaks
from oracle import emc_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("size,radius", [((23, 23, 23), (4.0, 4.0, 4.0)), ((31, 29, 27), (6.0, 5.0, 4.0))])
def test_sphere_mask(ctx, size, radius):
    m = BH_mask3d("sphere", size, radius, [0, 0, 0]).get()
    assert np.max(np.abs(m - O.mask3d_sphere(size, radius))) < 1e-5


def _planted(rng, shape=(40, 44, 48)):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    mag = 0.05 * rng.standard_normal(shape)
    for (cx, cy, cz), a in zip(((20, 20, 22), (30, 12, 15), (12, 30, 33), (34, 34, 8)), (5.0, 4.0, 3.0, 2.5)):
        mag += a * np.exp(-((x - cx - 0.3) ** 2 + (y - cy + 0.2) ** 2 + (z - cz) ** 2) / 2.0)
    ang = rng.integers(1, 20, size=shape).astype(np.int32)
    ang[0, 0, 0] = 0
    return mag.astype(np.float32), ang


def _compare(pk, rk):
    assert pk.shape == rk.shape, (pk, rk)
    assert np.max(np.abs(pk[:, :3] - rk[:, :3])) < 1e-4
    assert np.array_equal(pk[:, 3:6], rk[:, 3:6])
    assert np.max(np.abs(pk[:, 6] - rk[:, 6]) / np.abs(rk[:, 6])) < 1e-5


def test_peak_loop_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    mag, ang = _planted(rng)
    rm = 11
    grid = np.stack([np.arange(19) * 10.0, np.full(19, 5.0), np.arange(19) * -3.0], 1)
    mask = O.mask3d_sphere((2 * rm + 1,) * 3, (4.0, 4.0, 4.0))
    pm, pa = np.pad(mag, rm), np.pad(ang, rm)
    rk = O.tsearch_peaks(pm.astype(np.float64), pa, grid, mask, rm, 0.98, 1.0, 10, 1000)
    d_mag = DeviceArray.from_numpy(pm)
    d_ang = DeviceArray.from_numpy(pa)
    d_mask = DeviceArray.from_numpy(mask.astype(np.float32))
    out = np.zeros((10, 7))
    nf = _lib.C.c_int()
    pd = _lib.C.POINTER(_lib.C.c_double)
    _lib.call("emc_tsearch_peaks", ctx.handle, d_mag.vp, d_ang.vp, _lib.i3(d_mag.dims), d_mask.vp, rm,
              np.ascontiguousarray(grid).ctypes.data_as(pd), grid.shape[0], 0, 0.98, 1.0, 10, 1000,
              out.ctypes.data_as(pd), _lib.C.byref(nf))
    assert nf.value == 4
    _compare(out[:nf.value], rk)


def test_extract_peaks_driver(ctx):
    """extract_peaks (crop by szK, pad by szK + rmDim / rmDim, erase-mask
    cut-off search, loop, sampling) against the same steps on the oracle."""
    rng = np.random.default_rng(5)
    mag, ang = _planted(rng)
    grid = np.stack([np.arange(19) * 10.0, np.full(19, 5.0), np.zeros(19)], 1)
    szK, er = np.array([4, 4, 4]), np.array([4.0, 4.0, 4.0])
    pk = extract_peaks(mag, ang, grid, szK, er, peak_threshold=5, high_thr=1.0, sampling_rate=2)
    rm = int(max(er.max(), szK.max())) + 7
    m = mag.astype(np.float64)[4:-4, 4:-4, 4:-4]
    m = np.pad(m, 4 + rm)
    a = np.pad(ang, rm)
    mask = O.mask3d_sphere((2 * rm + 1,) * 3, er)
    cut = 0.98
    while np.sum(mask > cut) < 4.0 / 3.0 * np.pi * np.prod(er):
        cut *= 0.99
    rk = O.tsearch_peaks(m, a, grid, mask, rm, cut, 1.0, 10, 10000)
    rk[:, :3] *= 2
    _compare(pk, rk)


def _multi(ctx, pm, pa, grid, mask, rm, cutoff, thr, max_peaks, max_tries, n_peaks, sampling, normals):
    d_mag = DeviceArray.from_numpy(pm.astype(np.float32))
    d_ang = DeviceArray.from_numpy(pa.astype(np.int32))
    d_mask = DeviceArray.from_numpy(mask.astype(np.float32))
    rows = np.zeros((max_peaks, 10 * n_peaks))
    nf = _lib.C.c_int()
    pd = _lib.C.POINTER(_lib.C.c_double)
    nrm = np.ascontiguousarray(normals, dtype=np.float64)
    _lib.call("emc_tsearch_peaks_multi", ctx.handle, d_mag.vp, d_ang.vp, _lib.i3(d_mag.dims), d_mask.vp, rm,
              np.ascontiguousarray(grid).ctypes.data_as(pd), grid.shape[0], 0, cutoff, thr, max_peaks, max_tries,
              n_peaks, sampling, nrm.ctypes.data_as(pd), rows.ctypes.data_as(pd), _lib.C.byref(nf))
    return rows[:nf.value]


def _compare_rows(pm, rm_, n_peaks):
    """positions (and the random angles built on them) 1e-4, id-map angles
    exact, scores 1e-5 relative, peakMat's unused columns zero."""
    assert pm.shape == rm_.shape, (pm.shape, rm_.shape)
    a, b = pm.reshape(-1, n_peaks, 10), rm_.reshape(-1, n_peaks, 10)
    assert np.max(np.abs(a[:, :, 0:6] - b[:, :, 0:6])) < 1e-4
    assert np.array_equal(a[:, 0, 3:6], b[:, 0, 3:6])
    assert np.max(np.abs(a[:, :, 9] - b[:, :, 9]) / np.abs(b[:, :, 9])) < 1e-5
    assert not np.any(a[:, :, 6:9])


@pytest.mark.parametrize("n_peaks", [1, 3, 8])
def test_peak_loop_extra_peaks_match_oracle(ctx, n_peaks):
    rng = np.random.default_rng(11)
    mag, ang = _planted(rng)
    rm = 11
    grid = np.stack([np.arange(19) * 10.0, np.full(19, 5.0), np.arange(19) * -3.0], 1)
    mask = O.mask3d_sphere((2 * rm + 1,) * 3, (4.0, 4.0, 4.0))
    pm, pa = np.pad(mag, rm), np.pad(ang, rm)
    normals = rng.standard_normal((10, max(n_peaks - 1, 1), 3))
    rk = O.tsearch_peaks_multi(pm.astype(np.float64), pa, grid, mask, rm, 0.98, 1.0, 10, 1000, n_peaks, 2.5,
                               normals)
    pk = _multi(ctx, pm, pa, grid, mask, rm, 0.98, 1.0, 10, 1000, n_peaks, 2.5, normals)
    assert len(pk) == 4
    _compare_rows(pk, rk, n_peaks)


@pytest.mark.parametrize("head_id,n_peaks", [(0, 4), (6, 3)])
def test_extra_peaks_random_and_from_the_id_map(ctx, head_id, n_peaks):
    """The two branches of the extra peaks on a hand-made box (one non-zero
    voxel of another id): random angles when the id map's head is 0, the
    id map's angles and the class mean when it is positive."""
    mag = np.zeros((22, 22, 22), np.float32)
    ang = np.full((22, 22, 22), 5, np.int32)
    ang.reshape(-1)[:343] = head_id
    mag[12, 10, 11], mag[12, 10, 13], mag[12, 9, 11] = 4.0, 1.0, 0.5
    ang[12, 9, 11] = 7
    grid = np.stack([np.arange(9) * 10.0, np.full(9, 5.0), np.arange(9) * -2.0], 1)
    mask = O.mask3d_sphere((7, 7, 7), (2.0, 2.0, 2.0))
    normals = np.array([[[0.5, -1.0, 2.0], [1.5, 0.25, -0.5], [-2.0, 1.0, 0.125]]])[:, :n_peaks - 1]
    rk = O.tsearch_peaks_multi(mag.astype(np.float64), ang, grid, mask, 3, 0.5, 1.0, 1, 100, n_peaks, 2.0, normals)
    pk = _multi(ctx, mag, ang, grid, mask, 3, 0.5, 1.0, 1, 100, n_peaks, 2.0, normals)
    _compare_rows(pk, rk, n_peaks)
    if head_id:
        assert np.array_equal(pk[0, 13:16], grid[5]) and pk[0, 19] == 0.5


def test_extract_peaks_driver_extra_peaks(ctx):
    rng = np.random.default_rng(5)
    mag, ang = _planted(rng)
    grid = np.stack([np.arange(19) * 10.0, np.full(19, 5.0), np.zeros(19)], 1)
    szK, er = np.array([4, 4, 4]), np.array([4.0, 4.0, 4.0])
    pk = extract_peaks(mag, ang, grid, szK, er, peak_threshold=5, high_thr=1.0, sampling_rate=2, n_peaks=3, rng=17)
    rm = int(max(er.max(), szK.max())) + 7
    m = np.pad(mag.astype(np.float64)[4:-4, 4:-4, 4:-4], 4 + rm)
    a = np.pad(ang, rm)
    mask = O.mask3d_sphere((2 * rm + 1,) * 3, er)
    cut = 0.98
    while np.sum(mask > cut) < 4.0 / 3.0 * np.pi * np.prod(er):
        cut *= 0.99
    normals = np.random.default_rng(17).standard_normal((10, 2, 3))
    rk = O.tsearch_peaks_multi(m, a, grid, mask, rm, cut, 1.0, 10, 10000, 3, 2.0, normals).reshape(-1, 3, 10)
    assert pk.shape == rk.shape[:2] + (7,)
    assert np.max(np.abs(pk[:, :, 0:6] - rk[:, :, 0:6])) < 1e-4
    assert np.max(np.abs(pk[:, :, 6] - rk[:, :, 9]) / np.abs(rk[:, :, 9])) < 1e-5
