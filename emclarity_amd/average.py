"""Subtomogram averaging on the device: the per-particle body of
transformations/BH_average3d.m:1417-1441 (inverse rotation with symmetry,
masked normalisation) and the half-set sums of :1465-1490, for particles cut
from a resident tomogram (emc_extract_subtomos without the taper, :1345) and
accumulated a batch at a time (emc_average_accumulate)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import Context, DeviceArray, as_device, default_context
from .geometry import symmetry_matrices
from .sizing import BH_isWindowValid
from .subtomo import ResidentTomogram, extract_subtomos


class HalfSetSums:
    """The running sums of one class: particles and wedges of the odd (half
    set 1) and even (2) halves, and the particle counts (nExtracted)."""

    def __init__(self, window, wedge_size=None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        w = [int(x) for x in np.broadcast_to(np.asarray(window), (3,))]
        self.window = tuple(w)
        shape = (w[2], w[1], w[0])
        self.odd = DeviceArray.zeros(shape, np.float32, self.ctx)
        self.eve = DeviceArray.zeros(shape, np.float32, self.ctx)
        self.wedge_size = None
        self.wedge_odd = self.wedge_eve = None
        if wedge_size is not None:
            c = [int(x) for x in np.broadcast_to(np.asarray(wedge_size), (3,))]
            self.wedge_size = tuple(c)
            self.wedge_odd = DeviceArray.zeros((c[2], c[1], c[0]), np.float32, self.ctx)
            self.wedge_eve = DeviceArray.zeros((c[2], c[1], c[0]), np.float32, self.ctx)
        self.n_extracted = np.zeros(2, np.int64)


def average_accumulate(sums: HalfSetSums, windows: DeviceArray, mats, shifts, half, weights, interp_mask,
                       wedges: DeviceArray | None = None, symmetry: str = "C1") -> np.ndarray:
    """Add a batch of windows (n, wz, wy, wx) to the half-set sums. mats:
    (n, 9) column-major geometry columns 17:25 or (n, 3, 3); shifts (n, 3);
    half (n,) in {1, 2}; weights (n,). Returns the used flags (False where the
    reference would find a non-finite particle and ignore it)."""
    n = windows.shape[0]
    m = np.asarray(mats, dtype=np.float64)
    m = m.reshape(n, 9) if m.ndim == 2 else np.ascontiguousarray(m.transpose(0, 2, 1)).reshape(n, 9)
    m = np.ascontiguousarray(m)
    sh = np.ascontiguousarray(np.asarray(shifts, dtype=np.float64).reshape(n, 3))
    hf = np.ascontiguousarray(np.asarray(half, dtype=np.int32).reshape(n))
    wt = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(n))
    S = symmetry_matrices(symmetry)
    nsym = len(S)
    s_cm = np.ascontiguousarray(np.stack([np.asarray(x).reshape(-1, order="F") for x in S])) if nsym > 1 else None
    mask = as_device(interp_mask, np.float32, sums.ctx)
    used = np.zeros(n, np.int32)
    pd = C.POINTER(C.c_double)
    wdims = _lib.i3(sums.wedge_size) if wedges is not None else None
    _lib.call("emc_average_accumulate", sums.ctx.handle, n, _lib.i3(sums.window), windows.vp, m.ctypes.data_as(pd),
              sh.ctypes.data_as(pd), hf.ctypes.data_as(C.POINTER(C.c_int32)), wt.ctypes.data_as(pd), nsym,
              None if s_cm is None else s_cm.ctypes.data_as(pd), mask.vp, wdims,
              None if wedges is None else wedges.vp, sums.odd.vp, sums.eve.vp,
              None if wedges is None else sums.wedge_odd.vp, None if wedges is None else sums.wedge_eve.vp,
              used.ctypes.data_as(C.POINTER(C.c_int32)))
    u = used.astype(bool)
    for h in (1, 2):
        sums.n_extracted[h - 1] += int(np.sum(u & (hf == h)))
    return u


def weight_ang_check_peaks(position_row, n_peaks, score_sigma, track_stats=False):
    """BH_weightAngCheckPeaks (logicals/BH_weightAngCheckPeaks.m:1-128):
    the weights of one particle's n_peaks peaks (26 columns each, class in
    column 26 of each, -9999 = ignored) and its row re-ordered. Returns
    (peak_wgt (n_peaks,), sorted_row); -9999 marks a peak to skip.

    One included peak: it goes first, the other columns follow in order,
    weights [1, -9999, ...] (:16-26). Otherwise the included peaks sorted
    by score, descending (stable), then the excluded ones (:28-46). A peak
    within 5 deg of another is dropped (:51-72); as the reference's loop is
    written (`for iComb = find(..)` iterates once, over a column, and the
    colon takes its first element), peak i is compared with peak i - 1 only.
    Then scores / score_sigma, peaks with |score| <= 0.1 dropped, weights
    (score / max)^c normalised to sum 1 in complex arithmetic (real part
    kept), c = 2 or sqrt(max score) with track_stats, and weights below 0.1
    dropped (:86-100)."""
    row = np.asarray(position_row, dtype=np.float64).reshape(-1)
    n = int(n_peaks)
    cls = row[25:26 * n:26]
    inc = np.flatnonzero(cls != -9999)
    exc = np.flatnonzero(cls == -9999)
    wgt = np.zeros(n)
    blk = lambda k: slice(26 * k, 26 * k + 26)
    out = np.zeros_like(row)
    if inc.size == 1:
        k = int(inc[0])
        out[0:26] = row[blk(k)]
        out[26:] = np.delete(row, np.arange(26 * k, 26 * k + 26))
        wgt[0], wgt[1:] = 1.0, -9999.0
        return wgt, out
    order = inc[np.argsort(-row[26 * inc], kind="stable")]
    for j, k in enumerate(list(order) + list(exc)):
        out[blk(j)] = row[blk(int(k))]
    if 26 * n < row.size:
        out[26 * n:] = row[26 * n:]
    n_inc = int(np.sum(out[25:26 * n:26] != -9999))
    wgt[:] = -9999.0
    wgt[:n_inc] = 1.0
    for i in range(n_inc - 1, 0, -1):        # peaks n_inc .. 2, 0-based i, each against i - 1
        r1 = out[26 * i + 16:26 * i + 25].reshape(3, 3, order="F")
        r2 = out[26 * (i - 1) + 16:26 * (i - 1) + 25].reshape(3, 3, order="F")
        c = (np.trace(r1 @ r2.T) - 1.0) / 2.0
        if np.degrees(np.arccos(np.clip(c, -1.0, 1.0))) < 5.0:
            out[26 * i + 25] = -9999
            wgt[i] = -9999.0
    score = out[0:26 * n:26] / score_sigma
    keep = (wgt != -9999) & (np.abs(score) > 0.1)
    score[~keep] = -9999.0
    wgt[~keep] = -9999.0
    if np.any(keep):
        cf = np.sqrt(np.max(score[keep]) + 0j) if track_stats else 2.0
        w = (score[keep] / np.max(score[keep]) + 0j) ** cf
        w = np.real(w / np.sum(w))
        wgt[keep] = w
        wgt[keep & (wgt < 0.1)] = -9999.0
    if np.any(np.isnan(wgt)):
        raise ValueError("Found a nan in the re-weighted peaks")
    return wgt, out


def average_tomogram(sums: HalfSetSums, tomo: ResidentTomogram, positionList, interp_mask, sampling=1,
                     bin_shift=(0, 0, 0), mask_radius=None, wedges=None, symmetry="C1", ccc_cutoff=0.0,
                     class_idx=0, batch=256, n_peaks=1, score_sigma=1.0, track_stats=False):
    """One tomogram of the averaging loop (BH_average3d.m:1180-1500):
    particles with any peak's |CCC| >= ccc_cutoff, any peak not ignored
    (class_idx 0) or the first peak of that class, both half sets; each
    peak cut at its positionList(:, 11:13) / sampling + bin_shift and added
    with shiftVAL as its shift and its weight. n_peaks > 1: each row is first
    weighted and re-ordered by weight_ang_check_peaks (:1212-1228, the
    re-ordered row returned), skipped peaks flagged -9999; a window the
    reference cannot use ('noUse') flags the row and skips its later peaks
    (:1497-1505); a non-finite particle flags the row (:1445-1452).
    ``wedges``: None, or a callable(row index) -> (c, c, c) wedge
    (emc_weight_mask per particle, :1248-1256) of sums.wedge_size, shared
    by the row's peaks."""
    pos = np.array(positionList, dtype=np.float64, copy=True)
    n_peaks = int(n_peaks)
    inc = (np.any(np.abs(pos[:, 0:26 * n_peaks:26]) >= ccc_cutoff, axis=1) & np.isin(pos[:, 6], (1, 2)))
    if class_idx:
        inc &= pos[:, 25] == class_idx
    else:
        inc &= np.any(pos[:, 25:26 * n_peaks:26] != -9999, axis=1)
    rows = np.flatnonzero(inc)
    # entries (row, peak, weight)
    ent_r, ent_k, ent_w = [], [], []
    for r in rows:
        if n_peaks > 1:
            wgt, pos[r] = weight_ang_check_peaks(pos[r], n_peaks, score_sigma, track_stats)
        else:
            wgt = np.ones(1)
        for k in range(n_peaks):
            if wgt[k] == -9999:
                pos[r, 26 * k + 25] = -9999
                continue
            ent_r.append(int(r))
            ent_k.append(k)
            ent_w.append(float(wgt[k]))
    ent_r, ent_k, ent_w = np.array(ent_r, dtype=np.int64), np.array(ent_k, dtype=np.int64), np.array(ent_w)
    w = sums.window
    mr = mask_radius if mask_radius is not None else 0.25 * min(w)
    # a 'noUse' window skips the row's later peaks: find them first, on the host
    if ent_r.size:
        centres_all = (pos[ent_r[:, None], 26 * ent_k[:, None] + np.arange(10, 13)] / sampling
                       + np.asarray(bin_shift, dtype=np.float64))
        ok = np.array([not isinstance(BH_isWindowValid(tomo.dims, w, mr, c)[0], str) for c in centres_all])
        dead = {}
        for e in np.flatnonzero(~ok):
            dead.setdefault(int(ent_r[e]), int(ent_k[e]))
        live = np.array([ent_r[e] not in dead or ent_k[e] < dead[int(ent_r[e])] for e in range(ent_r.size)], bool)
        for r in dead:
            pos[r, 25:26 * n_peaks:26] = -9999
        ent_r, ent_k, ent_w = ent_r[live], ent_k[live], ent_w[live]
    for b0 in range(0, ent_r.size, batch):
        rb, kb, wb = ent_r[b0:b0 + batch], ent_k[b0:b0 + batch], ent_w[b0:b0 + batch]
        cols = 26 * kb[:, None]
        centres = pos[rb[:, None], cols + np.arange(10, 13)] / sampling + np.asarray(bin_shift, dtype=np.float64)
        win, valid, shift_val = extract_subtomos(tomo, centres, w, mr, taper=False)
        pos[rb[~valid], 25] = -9999
        keep = np.flatnonzero(valid)
        if keep.size == 0:
            continue
        nv = w[0] * w[1] * w[2] * 4
        if keep.size == rb.size:
            batch_win = win
        else:
            batch_win = DeviceArray((keep.size, w[2], w[1], w[0]), np.float32, sums.ctx)
            for j, i in enumerate(keep):
                _lib.call("emc_memcpy_d2d", sums.ctx.handle, C.c_void_p(batch_win.ptr + j * nv),
                          C.c_void_p(win.ptr + int(i) * nv), C.c_size_t(nv))
        wd = None
        if wedges is not None:
            c = sums.wedge_size
            wd = DeviceArray((keep.size, c[2], c[1], c[0]), np.float32, sums.ctx)
            cb = c[0] * c[1] * c[2] * 4
            for j, i in enumerate(keep):
                one = as_device(wedges(int(rb[i])), np.float32, sums.ctx)
                _lib.call("emc_memcpy_d2d", sums.ctx.handle, C.c_void_p(wd.ptr + j * cb), one.vp, C.c_size_t(cb))
        used = average_accumulate(sums, batch_win, pos[rb[keep, None], cols[keep] + np.arange(16, 25)],
                                  shift_val[keep], pos[rb[keep], 6], wb[keep], interp_mask, wd, symmetry)
        for r in rb[keep][~used]:
            pos[r, 25:26 * n_peaks:26] = -9999
    return pos
