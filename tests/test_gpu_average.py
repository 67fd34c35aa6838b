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
