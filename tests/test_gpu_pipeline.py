"""Template search into batched refinement, end to end on a small tomogram:
a 320 x 320 x 192 tomogram of N(0, 1) noise holding 24 copies of a 64^3 blob
template at uniform random rotations and sub-voxel shifts (SNR 0.03 over
the copy's box), searched as one whole-tomogram chunk over the 10-degree
grid (15,244 orientations, the nearest grid point at most ~7.6 degrees from
a random rotation), the top peaks cut as 96^3 windows, masked by a soft sphere of
radius 26 (neighbouring copies reach into a window) and refined with the
adjoint from their search orientation.

Against the inserted copies: every copy found within 3 voxels by the search,
the search orientation within the grid's reach (under 9 degrees), and after
refinement every copy whose window lies inside the tomogram (centre at
least 39 voxels from every face: the window's mask radius, 0.4 x 96) within
1 degree and 0.5 voxel, one miss allowed."""
import numpy as np
import pytest

from emclarity_amd import synthetic as S
from emclarity_amd.geometry import eulerSearch
from emclarity_amd.masking import BH_bandpass3d
from emclarity_amd.pipeline import (refine_peaks, score_against_truth, search_peaks, synthetic_tomogram,
                                    whole_chunk)
from emclarity_amd.workload import local_matrix, render_fast

pytestmark = pytest.mark.gpu


def test_search_then_refine_recovers_inserted_copies(ctx):
    rng = np.random.default_rng(404)
    blobs = S.make_blobs(rng, 30, radius=0.3 * 64, sigma=(2.0, 4.5))
    tomo_size = (320, 320, 192)
    vol, inserted = synthetic_tomogram(405, tomo_size, blobs, 64, 24, 0.03)
    assert len(inserted) == 24
    assert whole_chunk(tomo_size, (64, 64, 64))[0] == [384, 384, 256]
    template = render_fast(blobs, 64).astype(np.float32)
    grid = eulerSearch("C1", 180, 10, 180, 10)
    peaks, _, _ = search_peaks(vol, template, grid, 2 * len(inserted), 22, ctx=ctx)
    starts = np.stack([local_matrix(p[3], p[4], p[5] + p[3]) for p in peaks])
    found = score_against_truth(inserted, starts, peaks[:, :3])
    print(f"search: {found}")
    assert found["recovered"] == len(inserted)
    assert found["angle_err_p95_deg"] < 9.0
    # the copies are dark (subtracted): the refinement's reference has the
    # particles' contrast
    reference = (-render_fast(blobs, 96)).astype(np.float32)
    bp = BH_bandpass3d((96, 96, 96), 1e-3, 400.0, 25.0, "GPU", 10.0)
    R, centres, cc, valid, back, _ = refine_peaks(vol, peaks, reference, bp, window=96, ctx=ctx, mask_radius=26)
    inner = [p for p in inserted if all(39 <= p.centre[k] <= tomo_size[k] - 40 for k in range(3))]
    assert len(inner) == 12
    refined = score_against_truth(inner, R, centres)
    print(f"refined ({len(inner)} copies inside): {refined}; judge fallbacks {int(back.sum())}")
    assert refined["recovered"] == len(inner)
    assert refined["within_1deg_0p5px"] >= len(inner) - 1, refined
