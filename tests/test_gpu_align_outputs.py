"""One tomogram's alignment pass end to end on the GPU: particles planted in
a tomogram at known orientations and sub-voxel positions, a geometry that is
off by up to 12 degrees and 2.5 voxels, then extraction from the resident
tomogram -> batched search -> device refinement -> alignResume rows ->
BH_rawAlignmentsApply. The updated geometry must hold every particle within
1 degree and 0.5 voxel of the truth; the 'noUse' and ignored particles follow
the reference's rules."""
import math

import numpy as np
import pytest

from emclarity_amd import synthetic as S
from emclarity_amd.alignment_outputs import (BH_rawAlignmentsApply, align_tomogram, angle_error_deg,
                                             read_align_resume, write_align_resume)
from emclarity_amd.subtomo import ResidentTomogram
from emclarity_amd.workload import local_grid, local_matrix, make_size_class, random_rotation

pytestmark = pytest.mark.gpu


def test_tomogram_pass_recovers_planted_particles(ctx, tmp_path):
    n = 64
    rng = np.random.default_rng(12)
    sc, blobs = make_size_class(n)
    nx, ny, nz = 320, 256, 96
    tomo = np.zeros((nz, ny, nx))
    geom, truth = [], []
    for k, (ix, iy) in enumerate([(x, y) for y in (48, 112, 176) for x in (48, 112, 176, 240)]):
        c_true = np.array([ix, iy, 48.0]) + rng.uniform(0, 1, 3)          # 1-based voxel coordinates
        A = random_rotation(rng)
        Rt = A @ local_matrix(rng.uniform(0, 360), rng.uniform(0.5, 10), rng.uniform(-10, 10))
        box = S.render(blobs, n, Rt, c_true - np.floor(c_true))
        lo = (np.floor(c_true).astype(int) - 1) - n // 2                   # box voxel n/2 at floor(c)
        tomo[lo[2]:lo[2] + n, lo[1]:lo[1] + n, lo[0]:lo[0] + n] += box
        d = rng.uniform(-2.5, 2.5, 3)
        row = np.zeros(26)
        row[1], row[3], row[4:9], row[25] = 2, k + 1, 1, 1
        row[10:13] = c_true - d
        row[16:25] = A.reshape(-1, order="F")
        geom.append(row)
        truth.append((Rt, c_true))
    sig = float(np.std(tomo[tomo != 0]))
    tomo += rng.normal(0.0, 0.7 * sig, tomo.shape)
    edge = geom[0].copy()
    edge[3], edge[10:13] = 90, (3.0, 3.0, 3.0)       # too close to the edge: 'noUse'
    ignored = geom[1].copy()
    ignored[3], ignored[25] = 91, -9999
    g = np.stack(geom + [edge, ignored])

    t = ResidentTomogram.from_numpy(tomo.astype(np.float32))
    rows, idx = align_tomogram(t, g, sc.ref, sc.bandpass, local_grid(), mask_radius=20)
    print("rows:", rows.shape, flush=True)
    assert rows.shape == (12, 10) and np.array_equal(idx, np.arange(12))
    p = tmp_path / "alignResume" / "run" / "tomo_1.txt"
    write_align_resume(p, rows)
    back = read_align_resume(p)
    assert np.allclose(back, rows, atol=5e-4)
    new = BH_rawAlignmentsApply({"tomo_1": g}, {"tomo_1": back}, 1, 1)["tomo_1"]
    ang = [angle_error_deg(new[i, 16:25].reshape(3, 3, order="F"), truth[i][0]) for i in range(12)]
    pos = [float(np.linalg.norm(new[i, 10:13] - truth[i][1])) for i in range(12)]
    print("angle errors:", np.round(ang, 3), "position errors:", np.round(pos, 3), flush=True)
    assert max(ang) <= 1.0 and max(pos) <= 0.5
    assert new[12, 25] == -9999 and np.array_equal(new[13], g[13])
    assert np.array_equal(new[:12, 0], back[:, 5])
