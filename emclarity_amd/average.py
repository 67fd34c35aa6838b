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


def average_tomogram(sums: HalfSetSums, tomo: ResidentTomogram, positionList, interp_mask, sampling=1,
                     bin_shift=(0, 0, 0), mask_radius=None, wedges=None, symmetry="C1", ccc_cutoff=0.0,
                     class_idx=0, batch=256):
    """One tomogram of the averaging loop (BH_average3d.m:1180-1500) for
    nPeaks = 1: particles with |CCC| >= ccc_cutoff, not ignored and (class_idx
    != 0) of that class, both half sets, cut at positionList(:, 11:13) /
    sampling + bin_shift and added with shiftVAL as their shift. Windows the
    reference cannot use ('noUse') and non-finite particles are flagged
    -9999 in the returned copy of positionList. ``wedges``: None, or a
    callable(row index) -> (c, c, c) wedge (emc_weight_mask per particle,
    :1248-1256) of sums.wedge_size."""
    pos = np.array(positionList, dtype=np.float64, copy=True)
    inc = (np.abs(pos[:, 0]) >= ccc_cutoff) & (pos[:, 25] != -9999) & np.isin(pos[:, 6], (1, 2))
    if class_idx:
        inc &= pos[:, 25] == class_idx
    rows = np.flatnonzero(inc)
    w = sums.window
    mr = mask_radius if mask_radius is not None else 0.25 * min(w)
    for b0 in range(0, rows.size, batch):
        rb = rows[b0:b0 + batch]
        centres = pos[rb, 10:13] / sampling + np.asarray(bin_shift, dtype=np.float64)
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
        used = average_accumulate(sums, batch_win, pos[rb[keep], 16:25], shift_val[keep], pos[rb[keep], 6],
                                  np.ones(keep.size), interp_mask, wd, symmetry)
        pos[rb[keep][~used], 25] = -9999
    return pos
