"""Template search followed by batched adjoint refinement of its peaks: the
template-matching path of BH_templateSearch3d_2 (search, greedy peak
extraction) feeding the subtomogram alignment of BH_alignRaw3d_v2 (window
cut, alignment) on one GPU.

* search: ``template_search`` over the tomogram. On MI355X the whole
  tomogram fits in HBM as one chunk. ``whole_chunk`` sizes that chunk to
  an FFT-friendly edge (factors 2, 3, 5) at least tomogram + template per
  axis. The reference's chunker (alignment/BH_templateSearch3d_2.m:339-351)
  cuts overlapping chunks to fit a smaller GPU; those overlaps recompute up
  to 8x the voxels.
* peaks: ``extract_peaks`` (greedy, erase mask), up to ``n_peaks``.
* windows: ``extract_subtomos`` cuts a box of ``window`` voxels around each
  peak (1-based centres, BH_isWindowValid), from the tomogram resident on
  the device.
* refinement: one ``BatchAligner`` over all windows. Each particle starts
  at its search orientation. The adjoint refinement runs two stages: 0.6 x
  the support first, then the full support. The search pose acts as
  fallback and judge (``batched.judge``).

``synthetic_tomogram`` builds the workload: Gaussian noise N(0, 1) with
copies of a blob template at uniform random orientations and sub-voxel
shifts, non-overlapping, scaled to a given SNR. ``score_against_truth``
reports recovery and pose errors against the inserted copies.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from .device import DeviceArray, default_context
from .workload import angle_error_deg, local_matrix, random_rotation, render_fast


def _smooth(n):
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += 1


def whole_chunk(tomo_size, template_size):
    """One chunk holding the whole tomogram: (size_chunk, valid_area,
    n_iters, tomo_pre, tomo_post) for ``template_search``. Each edge is
    the smallest 2,3,5-smooth even number >= tomogram + template edge. The
    valid area is the tomogram, centred by BH_multi_padVal's rule."""
    from .sizing import BH_multi_padVal
    chunk = []
    for t, k in zip(tomo_size, template_size):
        c = _smooth(int(t) + int(k))
        while c % 2:
            c = _smooth(c + 1)
        chunk.append(c)
    pad = BH_multi_padVal(list(tomo_size), chunk)
    pre = [int(v) for v in pad[0]]
    post = [int(v) for v in pad[1]]
    return chunk, [int(v) for v in tomo_size], [1, 1, 1], pre, post


@dataclass
class Inserted:
    centre: np.ndarray      # (x, y, z) true centre, 0-based voxel coordinates (sub-voxel)
    R: np.ndarray           # 3x3: the copy is the template rotated by R (render convention)


def synthetic_tomogram(seed, tomo_size, blobs, template_edge, count, snr):
    """(tomogram (z, y, x) float32, [Inserted]). Each copy is the blob
    template rendered at a uniform random rotation R and a sub-voxel shift
    U[-0.5, 0.5] inside a template_edge^3 box. The box sits at a random
    integer position, and no two boxes overlap. The copies are subtracted
    (dense = dark, as in a reconstructed tomogram). The amplitude makes the
    copy's variance over its box ``snr`` times the noise variance."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = (int(v) for v in tomo_size)
    tomo = rng.standard_normal((nz, ny, nx), dtype=np.float32)
    t0 = render_fast(blobs, template_edge)
    amp = math.sqrt(snr) / float(np.std(t0))
    h = template_edge // 2
    placed = []
    tries = 0
    while len(placed) < count and tries < 200 * count:
        tries += 1
        c = np.array([rng.integers(h, nx - h), rng.integers(h, ny - h), rng.integers(h, nz - h)])
        if any(np.max(np.abs(c - p.centre)) < template_edge for p in placed):
            continue
        R = random_rotation(rng)
        frac = rng.uniform(-0.5, 0.5, 3)
        copy = render_fast(blobs, template_edge, R, frac)
        x, y, z = c
        tomo[z - h:z - h + template_edge, y - h:y - h + template_edge, x - h:x - h + template_edge] -= \
            (amp * copy).astype(np.float32)
        placed.append(Inserted(c + frac, R))
    return tomo, placed


def search_peaks(tomo, template, grid, n_peaks, lattice_radius, ctx=None, chunking=None, progress=None):
    """Template search and greedy peak extraction. Returns (peaks (n, 7):
    1-based x, y, z, phi, theta, psi - phi, score; seconds of the search;
    seconds of the extraction). ``progress`` as for ``template_search``."""
    from .template_search import extract_peaks, template_search
    ctx = ctx or default_context()
    tsize = (tomo.shape[2], tomo.shape[1], tomo.shape[0])
    tdims = (template.shape[2], template.shape[1], template.shape[0])
    sc, va, it, pre, post = chunking or whole_chunk(tsize, tdims)
    t0 = time.perf_counter()
    mag, ang = template_search(tomo, template, grid, sc, va, it, pre, post, ctx=ctx, progress=progress)
    ts = time.perf_counter() - t0
    t0 = time.perf_counter()
    r = [int(lattice_radius)] * 3
    # the peak list's angle rows are (phi, theta, psi - phi), as the
    # reference's ANGLE_LIST (BH_templateSearch3d_2.m:740)
    g = np.asarray(grid, dtype=np.float64).reshape(-1, 3)
    angle_list = np.column_stack([g[:, 0], g[:, 1], g[:, 2] - g[:, 0]])
    peaks = extract_peaks(mag, ang, angle_list, r, r, n_peaks / 2.0, 0.0, ctx=ctx)
    tp = time.perf_counter() - t0
    return peaks[:n_peaks], ts, tp


def refine_peaks(tomo, peaks, reference, bandpass, window=96, iters=16, coarse=(0.6, 8), table="c64", ctx=None,
                 mask_radius=None):
    """Cut a window around every peak and refine its pose with the adjoint,
    the search pose kept where its CC is higher. ``reference`` must have the
    particles' contrast (for ``synthetic_tomogram``, whose copies are dark,
    the negated template). ``mask_radius``: a soft sphere (EMC_maskShape with
    its default roll-off) applied to every window first, so that neighbouring
    particles inside the window stay out of the score. Returns (R (n, 3, 3),
    centres (n, 3) 1-based sub-voxel, cc (n,), valid (n,) bool, fell back
    (n,) bool, seconds)."""
    from .batched import BatchAligner, judge
    from .subtomo import ResidentTomogram, extract_subtomos
    ctx = ctx or default_context()
    t0 = time.perf_counter()
    P = len(peaks)
    res = ResidentTomogram.from_numpy(tomo, ctx)
    centres = np.rint(peaks[:, :3])
    wins, valid, _ = extract_subtomos(res, centres, (window,) * 3, 0.4 * window, taper=True)
    nb = 4 * window ** 3
    views = [wins.view((window,) * 3, np.float32, nb * i) for i in range(P)]
    starts = np.stack([local_matrix(p[3], p[4], p[5] + p[3]) for p in peaks])
    al = BatchAligner(window, reference, bandpass, peak_radius=6, table=table, ctx=ctx)
    mask = None
    if mask_radius:
        from .masking import EMC_maskShape
        mask = EMC_maskShape("sphere", (window,) * 3, [float(mask_radius)] * 3, "gpu", {}, ctx=ctx)
    al.set_particles(views, [None] * P, starts, mask=mask)
    R0 = DeviceArray.from_numpy(np.ascontiguousarray(starts.reshape(P, 9)), ctx=ctx)
    s0 = DeviceArray.zeros((P, 3), np.float64, ctx)
    R = R0.copy()
    s = s0.copy()
    k = min(int(coarse[1]), iters) if coarse else 0
    if k:
        al.refine_device(R, s, iters=k, support_radius=coarse[0] * al.support)
    cc = al.refine_device(R, s, iters=iters - k)
    Rh, sh, cch, back = judge(al, starts, np.zeros((P, 3)), R.get().reshape(P, 3, 3), s.get(), cc.get())
    dt = time.perf_counter() - t0
    # the window's centre voxel (index window // 2) is the 1-based centre
    # floor(c); a particle shifted by s sits at centre + s
    return Rh, centres + sh, cch, valid, back, dt


def score_against_truth(inserted, R, centres, radius=3.0):
    """Match every inserted copy to the nearest found particle within
    ``radius`` voxels. Returns dict(recovered, of, orientation and shift
    errors of the matched particles: median, 95th percentile, within
    1 deg / 0.5 px). centres are 1-based; truth is 0-based."""
    if len(centres) == 0:
        return {"recovered": 0, "of": len(inserted)}
    c0 = np.asarray(centres) - 1.0
    ang, sh = [], []
    for p in inserted:
        d = np.max(np.abs(c0 - p.centre), axis=1)
        j = int(np.argmin(d))
        if d[j] <= radius:
            ang.append(angle_error_deg(R[j], p.R))
            sh.append(float(np.max(np.abs(c0[j] - p.centre))))
    ang, sh = np.array(ang), np.array(sh)
    out = {"recovered": int(len(ang)), "of": len(inserted), "recovery": round(len(ang) / len(inserted), 4)}
    if len(ang):
        ok = (ang <= 1.0) & (sh <= 0.5)
        out.update({"angle_err_median_deg": round(float(np.median(ang)), 3),
                    "angle_err_p95_deg": round(float(np.percentile(ang, 95)), 3),
                    "shift_err_median_px": round(float(np.median(sh)), 3),
                    "shift_err_p95_px": round(float(np.percentile(sh, 95)), 3),
                    "within_1deg_0p5px": int(ok.sum())})
    return out
