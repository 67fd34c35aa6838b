"""Template search chunk loop of BH_templateSearch3d_2 on the GPU.

``template_search`` follows alignment/BH_templateSearch3d_2.m:397-398
(symmetric padding of the tomogram), :480-535 (chunk cut-out), :650-982
(theta -> chunk -> phi -> psi loop with the per-voxel max/argmax update) and
:985-994 (crop of the padding). The per-chunk work runs in
emc_template_search_chunk. ``preprocess_chunks`` is the reference's chunk
pre-processing (:480-611: standardise, bandpass, local normalisation over
the stats radius, sign and valid-area mask, global variance), one
emc_tsearch_preprocess_chunk call per chunk; ``template_search`` runs it
first when given its parameters, else it takes the tomogram as already
pre-processed.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceArray, as_device
from .sizing import BH_multi_padVal


def chunk_geometry(size_tomo, size_chunk, valid_area, n_iters, tomo_pre, tomo_post, size_template):
    """Chunk corners (0-based, padded tomogram), pad of the template into the
    chunk (padBIN, :372) and the valid-area offset inside a chunk (vA, :403)."""
    pad_bin = BH_multi_padVal(size_template, size_chunk)
    vA = BH_multi_padVal(valid_area, size_chunk)
    cuts = []
    for ix in range(int(n_iters[0])):
        for iy in range(int(n_iters[1])):
            for iz in range(int(n_iters[2])):
                cuts.append((ix * int(valid_area[0]), iy * int(valid_area[1]), iz * int(valid_area[2])))
    return cuts, pad_bin[0], vA[0]


def symmetric_pad(vol, pre, post):
    """padarray(..., 'symmetric') pre and post on (z,y,x) arrays, sizes (x,y,z)."""
    p = [(int(pre[2]), int(post[2])), (int(pre[1]), int(post[1])), (int(pre[0]), int(post[0]))]
    return np.pad(vol, p, mode="symmetric")


def preprocess_chunks(padded, corners, size_chunk, valid_calc, bp_vals, pixel, stats_radius,
                      should_be_ctf=1, ctx=None):
    """BH_templateSearch3d_2.m:480-611 on the padded tomogram (numpy, z,y,x):
    each chunk (x,y,z corner in ``corners``) is standardised, band-passed
    (bp_vals = HIGH_THRESH, HIGH_CUT, LOW_CUT at ``pixel``), locally
    normalised over ``stats_radius`` voxels, multiplied by -shouldBeCTF and
    the valid-area mask BH_mask3d('rectangle', sizeChunk, validCalc/2)
    (:402). The reference accumulates the square of each chunk's masked sum
    (:590), so globalVariance = sum(s_i^2)/N - (sum(s_i)/N)^2 with N = number
    of chunks * numel(chunk) (:603); every chunk is divided by its square
    root (:605-611). Returns (device chunks, globalVariance)."""
    from .masking import BH_mask3d
    sc = [int(s) for s in size_chunk]
    mask = BH_mask3d("rectangle", sc, np.asarray(valid_calc, dtype=np.float64) / 2.0, [0, 0, 0], ctx=ctx)
    ctx = mask.ctx
    bp = np.ascontiguousarray([float(bp_vals[0]), float(bp_vals[1]), float(bp_vals[2]), float(pixel)])
    dp = _lib.C.POINTER(_lib.C.c_double)
    out, fx, fx2, fn = [], 0.0, 0.0, 0.0
    for (cx, cy, cz) in corners:
        sub = padded[cz:cz + sc[2], cy:cy + sc[1], cx:cx + sc[0]]
        if sub.shape != (sc[2], sc[1], sc[0]):
            raise ValueError("chunk outside the padded tomogram")
        raw = DeviceArray.from_numpy(np.ascontiguousarray(sub, dtype=np.float32), ctx=ctx)
        res = DeviceArray(raw.shape, np.float32, ctx)
        tmp = _lib.C.c_double()
        _lib.call("emc_tsearch_preprocess_chunk", ctx.handle, raw.vp, _lib.i3(sc), bp.ctypes.data_as(dp),
                  int(stats_radius), float(-should_be_ctf), mask.vp, res.vp, _lib.C.byref(tmp))
        fx += tmp.value
        fx2 += tmp.value ** 2
        fn += float(np.prod(sc))
        out.append(res)
    gvar = fx2 / fn - (fx / fn) ** 2
    for c in out:
        _lib.call("emc_scale", ctx.handle, c.vp, int(np.prod(sc)), 1.0 / np.sqrt(gvar))
    return out, gvar


def template_search(tomogram, template, angles_pts, size_chunk, valid_area, n_iters, tomo_pre,
                    tomo_post, ctx=None, preprocess=None):
    """Per-voxel max CC and angle id (1-based into angles_pts, 0 where no
    orientation scored above 0 in a chunk other than the first, as the
    reference initialises RESULTS with zeros). Returns (peak, angle) with the
    tomogram's shape. ``angles_pts`` rows (phi, theta, psi) in the reference
    order theta -> phi -> psi; blocks of one theta are passed per chunk call.
    ``preprocess``: None (the tomogram is already pre-processed) or a dict
    with valid_calc, bp_vals, pixel, stats_radius and should_be_ctf for
    ``preprocess_chunks``."""
    tomogram = np.asarray(tomogram, dtype=np.float32)
    template = np.asarray(template, dtype=np.float32)
    nz, ny, nx = tomogram.shape
    padded = symmetric_pad(tomogram, tomo_pre, tomo_post)
    pnz, pny, pnx = padded.shape
    sc = [int(s) for s in size_chunk]
    va = [int(s) for s in valid_area]
    td = (template.shape[2], template.shape[1], template.shape[0])
    cuts, pad_low, v_low = chunk_geometry((pnx, pny, pnz), sc, va, n_iters, tomo_pre, tomo_post, td)
    t_dev = as_device(template, np.float32, ctx)
    ctx = t_dev.ctx
    pts = np.ascontiguousarray(np.asarray(angles_pts, dtype=np.float64).reshape(-1, 3))
    # group the flat list by theta, the outer loop of the reference (:650)
    thetas = []
    for i, th in enumerate(pts[:, 1]):
        if not thetas or thetas[-1][0] != th:
            thetas.append([th, i, i + 1])
        else:
            thetas[-1][2] = i + 1
    if preprocess is not None:
        chunks, _ = preprocess_chunks(padded, cuts, sc, ctx=ctx, **preprocess)
    else:
        chunks = []
        for (cx, cy, cz) in cuts:
            sub = padded[cz:cz + sc[2], cy:cy + sc[1], cx:cx + sc[0]]
            if sub.shape != (sc[2], sc[1], sc[0]):
                raise ValueError("chunk outside the padded tomogram")
            chunks.append(DeviceArray.from_numpy(np.ascontiguousarray(sub), ctx=ctx))
    vshape = (va[2], va[1], va[0])
    peaks = [DeviceArray.zeros(vshape, np.float32, ctx) for _ in cuts]
    angs = [DeviceArray.zeros(vshape, np.int32, ctx) for _ in cuts]
    dp = _lib.C.POINTER(_lib.C.c_double)
    first = True
    for (_, a0, a1) in thetas:
        for k, ch in enumerate(chunks):
            _lib.call("emc_template_search_chunk", ctx.handle, ch.vp, _lib.i3(sc), t_dev.vp,
                      _lib.i3(td), _lib.i3(pad_low), _lib.i3(v_low), _lib.i3(va),
                      pts[a0:a1].ctypes.data_as(dp), int(a1 - a0), int(a0 + 1), int(first and k == 0),
                      peaks[k].vp, angs[k].vp)
        first = False
    res_peak = np.zeros(padded.shape, np.float32)
    res_ang = np.zeros(padded.shape, np.int32)
    for (cx, cy, cz), p, a in zip(cuts, peaks, angs):
        z0, y0, x0 = cz + v_low[2], cy + v_low[1], cx + v_low[0]
        res_peak[z0:z0 + va[2], y0:y0 + va[1], x0:x0 + va[0]] = p.get()
        res_ang[z0:z0 + va[2], y0:y0 + va[1], x0:x0 + va[0]] = a.get()
    pre = [int(v) for v in tomo_pre]
    res_peak = res_peak[pre[2]:pre[2] + nz, pre[1]:pre[1] + ny, pre[0]:pre[0] + nx]
    res_ang = res_ang[pre[2]:pre[2] + nz, pre[1]:pre[1] + ny, pre[0]:pre[0] + nx]
    return res_peak, res_ang
