"""GPU parity of the MRC upload and the batched subtomogram extraction
against the FP64 oracle: uploads and type conversion bit-exact, windows
within 1e-6 of the tomogram's maximum (the FP32 taper weights), the untapered
interior bit-exact, 'noUse' flags and shiftVAL exact."""
import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.subtomo import ResidentTomogram, TomogramCache, extract_subtomos
from oracle import emc_oracle as O

from test_mrc import _data, _mrc_bytes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1, 2, 6])
@pytest.mark.parametrize("bo", ["<", ">"])
def test_upload_and_convert(ctx, tmp_path, mode, bo):
    d = _data(mode, shape=(12, 33, 35), seed=20 + mode)
    p = tmp_path / "t.mrc"
    p.write_bytes(_mrc_bytes(d, mode, bo, next_bytes=32))
    t = ResidentTomogram.from_file(p)
    assert t.dims == (35, 33, 12) and t.mode == mode
    assert np.array_equal(t.to_float().get(), O.mrc_get_volume(p).astype(np.float32))
    part = ResidentTomogram.from_file(p, kRange=(4, 9))
    assert np.array_equal(part.to_float().get(), O.mrc_get_volume(p, k_range=(4, 9)).astype(np.float32))


def _centres(rng, dims, n):
    nx, ny, nz = dims
    c = np.stack([rng.uniform(1, nx, n), rng.uniform(1, ny, n), rng.uniform(1, nz, n)], 1)
    c[0] = [nx / 2 + 0.25, ny / 2 + 0.5, nz / 2 + 0.75]     # interior
    c[1] = [6.0, ny / 2, nz / 2]                              # padded low in x
    c[2] = [nx - 3.4, ny - 2.2, nz / 2]                       # padded high in x and y
    c[3] = [2.0, 2.0, 2.0]                                    # too little kept: 'noUse'
    return c


@pytest.mark.parametrize("mode,taper", [(1, True), (1, False), (2, True), (2, False), (6, True)])
@pytest.mark.parametrize("window", [(32, 32, 32), (27, 30, 25)])
def test_extract_matches_oracle(ctx, mode, taper, window):
    """2-byte modes read each quad of voxels with two aligned 8-byte loads
    (any alignment of the window's first voxel; one load per voxel where
    those would leave the tomogram, at the padded edges)."""
    rng = np.random.default_rng(sum(window) + mode)
    shape = (70, 80, 90)
    tomo = _data(mode, shape=shape, seed=mode).astype({1: np.int16, 2: np.float32, 6: np.uint16}[mode])
    t = ResidentTomogram.from_numpy(tomo)
    assert t.mode == mode
    mr = (4.0, 4.0, 4.0)
    c = _centres(rng, t.dims, 40)
    out, valid, shifts = extract_subtomos(t, c, window, mr, taper=taper)
    got = out.get()
    scale = float(np.max(np.abs(tomo)))
    assert not valid[3]
    for i in range(c.shape[0]):
        ref, sh = O.extract_subtomo(tomo, window, mr, c[i], taper=taper)
        assert np.allclose(shifts[i], sh, atol=0, rtol=0)
        if ref is None:
            assert not valid[i] and np.all(got[i] == 0)
            continue
        assert valid[i]
        assert np.max(np.abs(got[i] - ref)) <= 1e-6 * scale, i
        # away from every edge of the kept block the window is a plain copy
        ind, pads, _ = O.window_valid((90, 80, 70), window, mr, c[i])
        lo = pads[0] + 7
        hi = np.array(window) - pads[1] - 7
        if np.all(hi > lo):
            sl = (slice(lo[2], hi[2]), slice(lo[1], hi[1]), slice(lo[0], hi[0]))
            assert np.array_equal(got[i][sl], ref[sl].astype(np.float32))


def test_extract_from_file_and_cache(ctx, tmp_path):
    d = _data(1, shape=(40, 48, 52), seed=7)
    p = tmp_path / "c.mrc"
    p.write_bytes(_mrc_bytes(d, 1, ">"))
    cache = TomogramCache(budget_bytes=300_000)      # one 199,680-byte int16 tomogram
    t = cache.get(p)
    assert cache.get(p) is t and len(cache) == 1
    c = np.array([[26.5, 24.0, 20.25], [10.0, 30.0, 12.0]])
    out, valid, _ = extract_subtomos(t, c, (24, 24, 24), (3, 3, 3))
    for i in range(2):
        ref, _ = O.extract_subtomo(O.mrc_get_volume(p), (24, 24, 24), (3, 3, 3), c[i])
        assert valid[i] and np.max(np.abs(out.get()[i] - ref)) <= 1e-6 * 32768
    # a second file over the budget evicts the first
    q = tmp_path / "d.mrc"
    q.write_bytes(_mrc_bytes(d, 1, "<"))
    cache.get(q)
    assert len(cache) == 1
    cache.clear()


def test_extract_rejects_bad_input(ctx):
    t = ResidentTomogram.from_numpy(np.zeros((20, 20, 20), np.float32))
    with pytest.raises(_lib.EmcError):
        extract_subtomos(t, [[np.nan, 10, 10]], (16, 16, 16), (2, 2, 2))
    out, valid, _ = extract_subtomos(t, np.zeros((0, 3)), (16, 16, 16), (2, 2, 2))
    assert valid.size == 0
