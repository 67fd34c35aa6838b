"""Seeded inputs of the template-search workload: a 512 x 512 x 256 FP32
tomogram of Gaussian noise N(0, 1) with copies of a 48^3 template (30
Gaussian blobs) inserted at non-overlapping random positions, each at one of
a few orientations of the 15-degree search grid, and the chunk geometry the
reference's 'convolution' chunker gives it (480^3 chunks, [2, 2, 1]
iterations, BH_templateSearch3d_2.m:339-351 -> BH_multi_iterator.m:86-156).
Shared by the GPU parity tests (host numpy; no device work here)."""
from __future__ import annotations

import numpy as np

from emclarity_amd import synthetic as S
from emclarity_amd.geometry import eulerSearch
from emclarity_amd.sizing import BH_multi_padVal, iterator_convolution
from oracle import emc_oracle as O

TOMO = (512, 512, 256)      # (x, y, z)
TEMPLATE = 48
LATTICE = 18                # lattice radius (px): particle size 2 * 18


def grid_15():
    """eulerSearch [180, 15, 180, 15]: 4,600 orientations."""
    return eulerSearch("C1", 180, 15, 180, 15)


def template(seed=41):
    rng = np.random.default_rng(seed)
    blobs = S.make_blobs(rng, 30, radius=0.3 * TEMPLATE, sigma=(1.5, 3.5))
    return S.render(blobs, TEMPLATE).astype(np.float32)


def tomogram(seed, orientations, count=150, amplitude=0.5):
    """(tomogram (z, y, x) float32, template, [(x, y, z, orientation row)]).
    The copies are inserted with negative contrast (dense = dark, as in a
    reconstructed tomogram): the search's swapQuadrants phase is -1 for
    480-voxel chunks, so the pre-processing's sign flip (-shouldBeCTF,
    BH_templateSearch3d_2.m:586) and this one make their peaks positive."""
    rng = np.random.default_rng(seed)
    t = template()
    tomo = rng.standard_normal((TOMO[2], TOMO[1], TOMO[0]), dtype=np.float32)
    h = TEMPLATE // 2
    rots = [O.interp3d(t.astype(np.float64), [p, th, ps - p], [0, 0, 0], "Bah", "forward")
            for (p, th, ps) in orientations]
    scale = amplitude / float(np.std(t))
    placed, tries = [], 0
    while len(placed) < count and tries < 100000:
        tries += 1
        c = np.array([rng.integers(h, TOMO[k] - h) for k in range(3)])
        if any(np.max(np.abs(c - q[:3])) < TEMPLATE for q in placed):
            continue
        k = int(rng.integers(len(orientations)))
        x, y, z = c
        tomo[z - h:z + h, y - h:y + h, x - h:x + h] -= (scale * rots[k]).astype(np.float32)
        placed.append(np.array([x, y, z, k]))
    return tomo, t, placed


def chunking():
    """The reference chunker for this tomogram and template."""
    r = iterator_convolution([512] * 3, list(TOMO), [TEMPLATE] * 3, [2 * LATTICE] * 3)
    return r


def chunk_offsets(size_chunk, valid_area):
    """padBIN of the template into a chunk, the valid area's low corner, and
    the trim rows the oracle takes (BH_multi_padVal)."""
    pad = BH_multi_padVal([TEMPLATE] * 3, size_chunk)
    va = BH_multi_padVal(valid_area, size_chunk)
    trim = BH_multi_padVal(size_chunk, valid_area)
    return pad, va[0], trim
