"""GPU parity of the batched averaging (emc_average_accumulate) against the
FP64 oracle, particle by particle: half-set sums within 2e-5 of their
maximum, C1 and C3, with wedges of their own size, a flat particle skipped;
two batches equal one; and one tomogram's averaging of planted particles
recovers the reference."""
import numpy as np
import pytest

from emclarity_amd import synthetic as S
from emclarity_amd.average import HalfSetSums, average_accumulate, average_tomogram
from emclarity_amd.device import DeviceArray
from emclarity_amd.subtomo import ResidentTomogram
from emclarity_amd.workload import random_rotation
from oracle import emc_oracle as O

from test_average import _mask

pytestmark = pytest.mark.gpu


def _batch(rng, n, w, c):
    wins = rng.standard_normal((n, w, w, w)).astype(np.float32) + 1.0
    mats = np.stack([random_rotation(rng).reshape(-1, order="F") for _ in range(n)])
    sh = rng.uniform(-1.5, 1.5, (n, 3))
    # flat, unrotated and unshifted: exactly constant inside the mask, rms 0, skipped
    wins[3], mats[3], sh[3] = 2.0, np.eye(3).reshape(-1), 0.0
    half = np.array([1, 2, 1, 1, 2, 2][:n])
    wt = rng.uniform(0.5, 1.5, n)
    wedges = rng.uniform(0, 1, (n, c, c, c)).astype(np.float32)
    return wins, mats, sh, half, wt, wedges


@pytest.mark.parametrize("symmetry", ["C1", "C3"])
def test_accumulate_matches_oracle(ctx, symmetry):
    rng = np.random.default_rng(5 if symmetry == "C1" else 6)
    n, w, c = 6, 32, 24
    wins, mats, sh, half, wt, wedges = _batch(rng, n, w, c)
    mask = _mask(w).astype(np.float32)
    sums = HalfSetSums(w, c)
    used = average_accumulate(sums, DeviceArray.from_numpy(wins), mats, sh, half, wt, mask,
                              DeviceArray.from_numpy(wedges), symmetry)
    odd, eve, wo, we, ref_used = O.average_particles(wins, mats, sh, half, wt, mask, wedges, symmetry)
    assert used.tolist() == ref_used.tolist() and not used[3]
    for got, ref in ((sums.odd, odd), (sums.eve, eve), (sums.wedge_odd, wo), (sums.wedge_eve, we)):
        g = got.get()
        assert np.max(np.abs(g - ref)) <= 2e-5 * np.max(np.abs(ref)), symmetry
    assert sums.n_extracted.tolist() == [2, 3]


def test_two_batches_equal_one(ctx):
    rng = np.random.default_rng(7)
    n, w, c = 6, 24, 16
    wins, mats, sh, half, wt, wedges = _batch(rng, n, w, c)
    mask = _mask(w).astype(np.float32)
    one, two = HalfSetSums(w, c), HalfSetSums(w, c)
    average_accumulate(one, DeviceArray.from_numpy(wins), mats, sh, half, wt, mask, DeviceArray.from_numpy(wedges))
    for sl in (slice(0, 2), slice(2, 6)):
        average_accumulate(two, DeviceArray.from_numpy(wins[sl]), mats[sl], sh[sl], half[sl], wt[sl], mask,
                           DeviceArray.from_numpy(wedges[sl]))
    for a, b in ((one.odd, two.odd), (one.eve, two.eve), (one.wedge_odd, two.wedge_odd)):
        assert np.array_equal(a.get(), b.get())


def test_tomogram_average_recovers_reference(ctx):
    n = 48
    rng = np.random.default_rng(8)
    blobs = S.make_blobs(rng, 30, radius=0.3 * n, sigma=(1.5, 3.5))
    ref = S.render(blobs, n)
    nx, ny, nz = 240, 192, 80
    tomo = np.zeros((nz, ny, nx))
    geom = []
    k = 0
    for iy in (40, 96, 152):
        for ix in (40, 96, 152, 200):
            c = np.array([ix, iy, 40.0]) + rng.uniform(0, 1, 3)
            R = random_rotation(rng)
            box = S.render(blobs, n, R, c - np.floor(c))
            lo = (np.floor(c).astype(int) - 1) - n // 2
            tomo[lo[2]:lo[2] + n, lo[1]:lo[1] + n, lo[0]:lo[0] + n] += box
            row = np.zeros(26)
            row[0], row[3], row[6], row[25] = 0.5, k + 1, 1 + k % 2, 1
            row[10:13] = c
            row[16:25] = R.reshape(-1, order="F")
            geom.append(row)
            k += 1
    edge = geom[0].copy()
    edge[3], edge[10:13] = 99, (2.0, 2.0, 2.0)
    geom = np.stack(geom + [edge])
    tomo += rng.normal(0, 0.5 * float(np.std(tomo[tomo != 0])), tomo.shape)
    mask = _mask(n).astype(np.float32)
    sums = HalfSetSums(n)
    out = average_tomogram(sums, ResidentTomogram.from_numpy(tomo.astype(np.float32)), geom, mask, mask_radius=12)
    assert out[-1, 25] == -9999 and np.all(out[:-1, 25] == 1)
    assert sums.n_extracted.tolist() == [6, 6]
    b = mask > 0.01
    for s in (sums.odd.get(), sums.eve.get()):
        cc = np.corrcoef(s[b], (ref * mask)[b])[0, 1]
        print("half-set CC with the reference:", round(float(cc), 4), flush=True)
        assert cc > 0.95


def test_tomogram_average_with_two_peaks(ctx):
    """nPeaks = 2 (BH_average3d.m:1212-1228 with BH_weightAngCheckPeaks):
    each particle's second peak 30 deg off the first, scores 0.8 / 0.5, so
    weights 1 / 1.390625 and 0.390625 / 1.390625; the sums equal one
    accumulation of the same peaks with those weights, bit for bit. A row
    whose better peak's window is 'noUse' is flagged and its other peak
    skipped."""
    from emclarity_amd.average import weight_ang_check_peaks
    from emclarity_amd.subtomo import extract_subtomos
    n = 32
    rng = np.random.default_rng(9)
    blobs = S.make_blobs(rng, 20, radius=0.3 * n, sigma=(1.5, 3.0))
    nx, ny, nz = 160, 128, 64
    tomo = np.zeros((nz, ny, nx))
    rows = []
    Rz = np.array([[np.cos(np.pi / 6), -np.sin(np.pi / 6), 0], [np.sin(np.pi / 6), np.cos(np.pi / 6), 0], [0, 0, 1]])
    for k, (ix, iy) in enumerate(((30, 30), (80, 30), (130, 30), (30, 90), (80, 90), (130, 90))):
        c = np.array([ix, iy, 32.0]) + rng.uniform(0, 1, 3)
        R = random_rotation(rng)
        box = S.render(blobs, n, R, c - np.floor(c))
        lo = (np.floor(c).astype(int) - 1) - n // 2
        tomo[lo[2]:lo[2] + n, lo[1]:lo[1] + n, lo[0]:lo[0] + n] += box
        row = np.zeros(52)
        for p, (score, M) in enumerate(((0.8, R), (0.5, R @ Rz))):
            row[26 * p], row[26 * p + 3], row[26 * p + 6], row[26 * p + 25] = score, k + 1, 1 + k % 2, 1
            row[26 * p + 10:26 * p + 13] = c
            row[26 * p + 16:26 * p + 25] = M.reshape(-1, order="F")
        rows.append(row)
    edge = rows[0].copy()
    edge[3], edge[10:13] = 99, (2.0, 2.0, 2.0)        # the 0.8 peak at the corner: 'noUse'
    geom = np.stack(rows + [edge])
    tomo += rng.normal(0, 0.3 * float(np.std(tomo[tomo != 0])), tomo.shape)
    rt = ResidentTomogram.from_numpy(tomo.astype(np.float32))
    mask = _mask(n).astype(np.float32)
    sums = HalfSetSums(n)
    out = average_tomogram(sums, rt, geom, mask, mask_radius=8, n_peaks=2)
    assert np.all(out[-1, 25::26] == -9999) and np.all(out[:-1, 25::26] == 1)
    assert sums.n_extracted.tolist() == [6, 6]
    wgt, _ = weight_ang_check_peaks(geom[0], 2, 1.0)
    assert np.allclose(wgt, [1 / 1.390625, 0.390625 / 1.390625], rtol=0, atol=1e-15)
    ref = HalfSetSums(n)
    cen = np.concatenate([g.reshape(2, 26)[:, 10:13] for g in geom[:-1]])
    win, valid, shift_val = extract_subtomos(rt, cen, (n, n, n), 8, taper=False)
    assert np.all(valid)
    mats = np.concatenate([g.reshape(2, 26)[:, 16:25] for g in geom[:-1]])
    half = np.repeat(geom[:-1, 6], 2)
    average_accumulate(ref, win, mats, shift_val, half, np.tile(wgt, 6), mask)
    for a, b in ((sums.odd, ref.odd), (sums.eve, ref.eve)):
        assert np.array_equal(a.get(), b.get())
