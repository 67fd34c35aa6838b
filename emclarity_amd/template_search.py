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
                      should_be_ctf=1, ctx=None, med_filt=0):
    """BH_templateSearch3d_2.m:480-611 on the padded tomogram (numpy, z,y,x):
    each chunk (x,y,z corner in ``corners``) is standardised, band-passed
    (bp_vals = HIGH_THRESH, HIGH_CUT, LOW_CUT at ``pixel``), locally
    median-filtered when ``med_filt`` is 3, 5 or 7 (Tmp_medianFilter,
    :492-499, :543-548), locally normalised over ``stats_radius`` voxels,
    multiplied by -shouldBeCTF and
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
                  int(med_filt), int(stats_radius), float(-should_be_ctf), mask.vp, res.vp, _lib.C.byref(tmp))
        fx += tmp.value
        fx2 += tmp.value ** 2
        fn += float(np.prod(sc))
        out.append(res)
    gvar = fx2 / fn - (fx / fn) ** 2
    for c in out:
        _lib.call("emc_scale", ctx.handle, c.vp, int(np.prod(sc)), 1.0 / np.sqrt(gvar))
    return out, gvar


def template_search(tomogram, template, angles_pts, size_chunk, valid_area, n_iters, tomo_pre,
                    tomo_post, ctx=None, preprocess=None, progress=None):
    """Per-voxel max CC and angle id (1-based into angles_pts, 0 where no
    orientation scored above 0 in a chunk other than the first, as the
    reference initialises RESULTS with zeros). Returns (peak, angle) with the
    tomogram's shape. ``angles_pts`` rows (phi, theta, psi) in the reference
    order theta -> phi -> psi; blocks of one theta are passed per chunk call.
    ``preprocess``: None (the tomogram is already pre-processed) or a dict
    with valid_calc, bp_vals, pixel, stats_radius and should_be_ctf for
    ``preprocess_chunks``. ``progress``: called as progress(done, total)
    after each theta block (all chunks)."""
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
    for ti, (_, a0, a1) in enumerate(thetas):
        for k, ch in enumerate(chunks):
            _lib.call("emc_template_search_chunk", ctx.handle, ch.vp, _lib.i3(sc), t_dev.vp,
                      _lib.i3(td), _lib.i3(pad_low), _lib.i3(v_low), _lib.i3(va),
                      pts[a0:a1].ctypes.data_as(dp), int(a1 - a0), int(a0 + 1), int(first and k == 0),
                      peaks[k].vp, angs[k].vp)
        first = False
        if progress is not None:
            ctx.synchronize()
            progress(ti + 1, len(thetas))
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


# ------------------------------------------------------------ peak extraction

def high_threshold(peak_threshold, tomo_size, n_angles_first):
    """highThr of BH_templateSearch3d_2.m:337: sqrt(2) erfcinv(ceil(0.1
    peakThreshold) 2 / (numel(tomogram) nAngles(1)))."""
    from scipy.special import erfcinv
    return float(np.sqrt(2.0) * erfcinv(np.ceil(peak_threshold * 0.10) * 2.0
                                        / (float(np.prod(tomo_size)) * float(n_angles_first))))


def extract_peaks(mag, ang, angle_list, lattice_radius, erase_radius, peak_threshold, high_thr,
                  sampling_rate=1, erase_type="sphere", max_tries=10000, convention="Bah", n_peaks=1, rng=None,
                  ctx=None):
    """Greedy peak extraction of BH_templateSearch3d_2.m:1036-1275 on the
    cropped score map ``mag`` and angle-id map ``ang`` ((z, y, x), ids 1-based
    into ``angle_list`` rows (phi, theta, psi-phi), 0 = none).
    lattice_radius (szK) and erase_radius are (x, y, z) in voxels. The map is
    cropped by szK and padded by szK + rmDim, the ids by rmDim, rmDim =
    max(erase, szK) + 7 (:1036-1080; the reference pads on the host, so do
    these copies); the erase mask is BH_mask3d(erase_type, 2 rmDim + 1,
    erase_radius) with its cut-off lowered by 1 % steps until it keeps the
    sphere's volume (:1102-1122); then emc_tsearch_peaks. Returns (n, 7):
    x, y, z (times sampling_rate), phi, theta, psi - phi, score.
    n_peaks > 1 (:1156-1240, emc_tsearch_peaks_multi): the centre of mass
    over the wider box, and per taken peak n_peaks entries, (n, n_peaks, 7),
    the extra ones with the taken position and their angles either from the
    id map or random (position + nRandom^2 x standard normals drawn from
    ``rng``, a numpy Generator or seed, where the reference calls randn);
    column 6 of an extra peak is its iSNR."""
    from .geometry import _CONV
    from .masking import BH_mask3d, BH_padZeros3d
    n_peaks = int(n_peaks)
    if n_peaks < 1:
        raise ValueError("n_peaks must be >= 1")
    szK = np.asarray(lattice_radius, dtype=np.int64).reshape(3)
    er = np.asarray(erase_radius, dtype=np.float64).reshape(3)
    rm = int(max(er.max(), szK.max())) + 7
    m = BH_padZeros3d(mag, -szK, -szK, ctx=ctx)
    m = BH_padZeros3d(m, szK + rm, szK + rm, ctx=ctx)
    ctx = m.ctx
    a = np.pad(np.asarray(ang, dtype=np.int32), rm)
    if a.shape != m.shape:
        raise ValueError("score and angle maps differ in size")
    d_ang = DeviceArray.from_numpy(a, ctx=ctx)
    e = 2 * rm + 1
    mask = BH_mask3d(erase_type, [e, e, e], er, [0, 0, 0], ctx=ctx)
    mh = mask.get()
    cutoff, tries = 0.98, 0
    pre = 0.0 if str(erase_type).lower() == "rectangle" else 4.0 / 3.0 * np.pi
    while np.sum(mh > cutoff) < pre * np.prod(er):
        cutoff *= 0.99
        tries += 1
        if tries > 1000:
            raise RuntimeError("Did not find an appropriate erase mask")
    al = np.ascontiguousarray(np.asarray(angle_list, dtype=np.float64).reshape(-1, 3))
    max_peaks = int(2 * peak_threshold)
    nf = _lib.C.c_int()
    pd = _lib.C.POINTER(_lib.C.c_double)
    if n_peaks > 1:
        normals = np.ascontiguousarray(np.random.default_rng(rng).standard_normal((max(max_peaks, 1), n_peaks - 1, 3)))
        rows = np.zeros((max(max_peaks, 1), 10 * n_peaks))
        _lib.call("emc_tsearch_peaks_multi", ctx.handle, m.vp, d_ang.vp, _lib.i3(m.dims), mask.vp, rm,
                  al.ctypes.data_as(pd), al.shape[0], _CONV[convention], float(cutoff), float(high_thr), max_peaks,
                  int(max_tries), n_peaks, float(sampling_rate), normals.ctypes.data_as(pd), rows.ctypes.data_as(pd),
                  _lib.C.byref(nf))
        r = rows[:nf.value].reshape(-1, n_peaks, 10)
        return np.concatenate([r[:, :, 0:6], r[:, :, 9:10]], axis=2)
    out = np.zeros((max(max_peaks, 1), 7))
    _lib.call("emc_tsearch_peaks", ctx.handle, m.vp, d_ang.vp, _lib.i3(m.dims), mask.vp, rm,
              al.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)), al.shape[0], _CONV[convention],
              float(cutoff), float(high_thr), max_peaks, int(max_tries),
              out.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)), _lib.C.byref(nf))
    peaks = out[:nf.value].copy()
    peaks[:, :3] *= sampling_rate
    return peaks


def _matlab_d(v):
    """MATLAB fprintf '%d' of a double: integers print as integers, other
    values switch to '%e'."""
    v = float(v)
    return str(int(v)) if v == int(v) else f"{v:e}"


def write_template_csv(csv_path, pos_path, peaks, sampling_rate, n_previous=0, convention="Bah",
                       symmetry_matrices=None):
    """The CSV and .pos files of BH_templateSearch3d_2.m:1283-1341 (consumed
    by BH_geometryInitialize): per peak i one row of 26 fields
    'score samplingRate 0 i+nPrevious 1 1 1 1 1 0 x y z phi theta psi r(9) 1'
    with r = BH_defineMatrix([phi, theta, psi], convention, 'inv') (times the
    in-plane symmetry matrix rem(n + S, S) + 1 when S > 1), column-major, and
    the .pos line x y z / samplingRate. Formats as the reference's fprintf.
    peaks (n, k, 7) (extract_peaks with n_peaks = k > 1): the k peaks of a
    row on one line, each with the first peak's score (:1310-1325)."""
    from .geometry import BH_defineMatrix
    peaks = np.asarray(peaks, dtype=np.float64)
    peaks = peaks.reshape(-1, 1, 7) if peaks.ndim < 3 else peaks
    S = len(symmetry_matrices) if symmetry_matrices is not None else 1
    with open(csv_path, "w") as fc, open(pos_path, "w") as fp:
        n = 1
        for i, pks in enumerate(peaks, start=1):
            if not np.all(pks[0, 0:3]):
                continue
            line = ""
            for pk in pks:
                r = BH_defineMatrix(pk[3:6], convention, "inv")
                if S > 1:
                    r = r @ np.asarray(symmetry_matrices[(n + S) % S])
                rv = r.reshape(-1, order="F")
                fields = ([f"{pks[0, 6]:1.2f}", _matlab_d(sampling_rate), "0", _matlab_d(i + n_previous)]
                          + ["1"] * 5 + ["0"] + [f"{v:f}" for v in pk[0:3]] + [_matlab_d(v) for v in pk[3:6]]
                          + [f"{v:f}" for v in rv] + ["1"])
                line += " ".join(fields) + " "
            fc.write(line + "\n")
            fp.write("%f %f %f\n" % tuple(pks[0, 0:3] / sampling_rate))
            n += 1
