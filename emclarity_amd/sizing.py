"""Integer sizing helpers of the path (A17): BH_multi_padVal,
BH_multi_iterator ('fourier', 'convolution'), BH_multi_validArea,
BH_isWindowValid."""
from __future__ import annotations

import math

import numpy as np

NEXT_BEST = [64, 72, 80, 84, 90, 96, 108, 112, 120, 126, 128, 144, 160, 162, 168, 180, 192, 216, 224,
             240, 256, 270, 288, 320, 324, 336, 360, 378, 384, 400, 416, 432, 448, 480, 486, 504, 512]
APODIZATION = 2 * 6


def BH_multi_padVal(size1, size2) -> np.ndarray:
    """coordinates/BH_multi_padVal.m:7-22 -> 4x3 [fwdLow; fwdTop; invLow; invTop]."""
    s1 = np.asarray(size1, dtype=np.int64)
    s2 = np.asarray(size2, dtype=np.int64)
    d = s2 - s1
    fwd = np.stack([d // 2 + (s1 % 2) * (d % 2), d // 2 + (1 - s1 % 2) * (d % 2)])
    d = s1 - s2
    inv = np.stack([d // 2 + (s2 % 2) * (d % 2), d // 2 + (1 - s2 % 2) * (d % 2)])
    return np.concatenate([fwd, inv])


def iterator_fourier(size, descend=False):
    """BH_multi_iterator(SIZE, 'fourier') (:86-110 of the 'fourier' case)."""
    out = []
    for s in size:
        if descend:
            c = [b for b in NEXT_BEST if b <= s]
            out.append(c[-1] if c else NEXT_BEST[0])
        else:
            c = [b for b in NEXT_BEST if b >= s]
            out.append(c[0] if c else NEXT_BEST[-1])
    return np.array(out)


def iterator_convolution(target, size_image, size_template, size_particle):
    """BH_multi_iterator([target; image; template; 2*latticeRadius], 'convolution')
    (coordinates/BH_multi_iterator.m, 'convolution' case). Returns a dict
    with borderSizeKeep (tomoPre), postPad (tomoPost), chunkSize, validAreaKeep,
    validAreaCalc and nIters."""
    target = np.asarray(target)
    if np.all(target == 256):
        nb = [128, 144, 160, 168, 192, 216, 224, 256]
    elif np.all(target == 384):
        nb = [128, 144, 160, 168, 192, 216, 224, 256, 288, 300, 320, 336, 360, 384]
    elif np.all(target == 432):
        nb = [128, 144, 160, 168, 192, 216, 224, 256, 288, 300, 320, 336, 360, 384, 400, 432]
    elif np.all(target == 512):
        nb = [128, 144, 160, 168, 192, 216, 224, 256, 288, 300, 320, 336, 360, 384, 400, 432, 480, 512]
    elif np.all(target > 512):
        nb = [128, 144, 160, 168, 192, 216, 224, 256, 288, 300, 320, 336, 360, 384, 400, 432, 480, 512,
              540, 576, 640, 648, 720, 756, 768, 810, 864, 896, 960, 972, 1008, 1024]
    else:
        raise ValueError("target size must be 256, 384, 432, 512 or > 512")
    img = np.asarray(size_image, dtype=np.int64)
    tmp = np.asarray(size_template, dtype=np.int64)
    part = np.asarray(size_particle, dtype=np.int64)
    border_calc = (tmp + APODIZATION) // 2
    border_keep = border_calc + 2 * part
    if abs(int(np.sum(img - tmp))) < np.sum(0.1 * img):
        return dict(tomoPre=np.zeros(3, int), tomoPost=np.zeros(3, int), chunkSize=img,
                    validAreaKeep=img, validAreaCalc=img, nIters=np.ones(3, int))
    nbv = np.asarray(nb, dtype=np.int64)[:, None]
    valid_calc = nbv - 2 * border_calc[None, :]
    valid_keep = nbv - 2 * border_keep[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        min_iter = np.floor(img[None, :] / valid_keep).astype(np.int64)
    post_pad = nbv - (img[None, :] + border_keep[None, :] - min_iter * (valid_keep + 1))
    with np.errstate(divide="ignore", invalid="ignore"):
        score = nbv / post_pad * (min_iter >= 0)
    score = np.where(np.isnan(score), -np.inf, score)      # MATLAB max ignores NaN
    c = [int(np.argmax(score[:, k])) for k in range(3)]   # MATLAB max: first maximum
    return dict(tomoPre=border_keep, tomoPost=np.array([post_pad[c[k], k] for k in range(3)]),
                chunkSize=np.array([nb[c[k]] for k in range(3)]),
                validAreaKeep=np.array([valid_keep[c[k], k] for k in range(3)]),
                validAreaCalc=np.array([valid_calc[c[k], k] for k in range(3)]),
                nIters=np.array([min_iter[c[k], k] + 1 for k in range(3)]))


def BH_multi_validArea(mask_size, mask_radius, scale_calc_size=1.5):
    """logicals/BH_multi_validArea.m:38-76 -> (sizeWindow, sizeCalc, sizeMask,
    padWindow, padCalc)."""
    size_mask = np.ceil(scale_calc_size * np.max(mask_size) * np.ones(3)).astype(np.int64)
    size_window = np.max(size_mask) * np.ones(3, dtype=np.int64)
    pad_window = BH_multi_padVal(size_mask, size_window)
    size_calc = iterator_fourier(scale_calc_size * 2 * np.max(mask_radius) * np.ones(3))
    if np.any(size_calc - size_window < 0):
        size_calc = iterator_fourier(size_window)
    pad_calc = BH_multi_padVal(size_mask, size_calc)
    return size_window, size_calc, size_mask, pad_window, pad_calc


def BH_isWindowValid(volume_size, window_size, mask_radius, center):
    """logicals/BH_isWindowValid.m:42-65 -> (indices 2x3 (1-based) or 'noUse',
    padvalues 2x3, shifts 3)."""
    vs = np.asarray(volume_size, dtype=np.float64)
    ws = np.asarray(window_size, dtype=np.float64)
    c = np.asarray(center, dtype=np.float64)
    min_size = (np.max(mask_radius) + 6) * 2 * np.ones(3)
    low_c = np.ceil((ws - 1) / 2)
    top_c = low_c - np.mod(ws - 1, 2)
    wc = np.floor(c)
    delta = c - wc
    low = wc - low_c
    top = vs - (wc + top_c)
    pad_low = (np.abs(low) + 1) * (low <= 0)
    pad_top = np.abs(top) * (top <= 0)
    ind = np.stack([low + pad_low, wc + top_c - pad_top])
    pad = np.stack([pad_low, pad_top])
    avail = ws - pad[0] - pad[1]
    if np.any(avail - min_size < -2):
        return "noUse", avail, delta
    return ind.astype(np.int64), pad.astype(np.int64), delta
