"""Alignment outputs: one tomogram's pass of the alignment loop on the batched
path, its result rows, the alignResume text files and the geometry update.

The reference aligns each particle of a tomogram at its current orientation
(rotation columns 17:25 of the geometry) and position (columns 11:13), and
writes one row per particle (alignment/BH_alignRaw3d_v2.m:1762-1765,
1862-1903):

    [class, particle id, phi, theta, psi - phi, CCC, 1, dx, dy, dz]

where the angles are the best relative rotation found by the local search
(defineMatrix(angles, 'Bah', 'inv') applied after the current orientation)
and (dx, dy, dz) the particle's displacement in binned pixels (the found
shift minus the windowing offset shiftVAL). BH_rawAlignmentsApply
(metaData/BH_rawAlignmentsApply.m:40-89) then composes the rotations and
moves the positions.
"""
from __future__ import annotations

import math
import os

import numpy as np

from .batched import BatchAligner
from .geometry import BH_defineMatrix
from .subtomo import ResidentTomogram, extract_subtomos

ROW_FORMAT = "%s %s %6.3f %6.3f %6.3f %6.6f %6.6f %6.3f %6.3f %6.3f "


def bah_angles_from_matrix(M) -> np.ndarray:
    """Angles (a1, a2, a3) in degrees with BH_defineMatrix([a1, a2, a3],
    'Bah', 'inv') = Rz(a1) Rx(a2) Rz(a3) = M, for one (3, 3) or a stack
    (..., 3, 3) of rotations; a2 in [0, 180]. Where sin(a2) = 0 only a1 + a3
    (a2 = 0) or a1 - a3 (a2 = 180) is fixed: a3 = 0 there."""
    M = np.asarray(M, dtype=np.float64)
    c = np.clip(M[..., 2, 2], -1.0, 1.0)
    a2 = np.arccos(c)
    s2 = np.sqrt(np.maximum(0.0, M[..., 0, 2] ** 2 + M[..., 1, 2] ** 2))
    gen = s2 > 1e-9
    a1 = np.where(gen, np.arctan2(M[..., 0, 2], -M[..., 1, 2]), np.arctan2(M[..., 1, 0], M[..., 0, 0]))
    a3 = np.where(gen, np.arctan2(M[..., 2, 0], M[..., 2, 1]), 0.0)
    return np.degrees(np.stack([a1, a2, a3], axis=-1))


def _matlab_d(v) -> str:
    v = float(v)
    return "%d" % int(v) if v == int(v) else "%e" % v


def format_align_resume(rows, n_peaks: int = 1) -> str:
    """The text of alignResume/<prefix>/<tomo>.txt (:1894-1903)."""
    rows = np.asarray(rows, dtype=np.float64).reshape(len(rows), 10 * n_peaks)
    lines = []
    for r in rows:
        line = ""
        for p in range(n_peaks):
            b = r[10 * p:10 * p + 10]
            line += ROW_FORMAT % ((_matlab_d(b[0]), _matlab_d(b[1])) + tuple(b[2:10]))
        lines.append(line + "\n")
    return "".join(lines)


def write_align_resume(path, rows, n_peaks: int = 1):
    os.makedirs(os.path.dirname(os.fspath(path)) or ".", exist_ok=True)
    with open(path, "w") as f:
        f.write(format_align_resume(rows, n_peaks))


def read_align_resume(path) -> np.ndarray:
    """MATLAB load() of a resume file (:784-790): whitespace-separated rows."""
    with open(path) as f:
        rows = [[float(t) for t in line.split()] for line in f if line.strip()]
    return np.array(rows, dtype=np.float64).reshape(len(rows), -1)


def unique_particle_rows(sort_ccc) -> np.ndarray:
    """:1866-1884: rows with particle id 0 dropped, then the first row of
    each particle id, in order."""
    s = np.asarray(sort_ccc, dtype=np.float64)
    s = s[s[:, 1] != 0]
    _, first = np.unique(s[:, 1], return_index=True)
    return s[np.sort(first)]


def revert_drops(best, initial, drop: float = 0.15) -> np.ndarray:
    """:1784-1792: a row whose CCC (column 6) fell below the initial one by
    more than ``drop`` of the initial CCC goes back to the initial row."""
    best = np.array(best, dtype=np.float64, copy=True)
    initial = np.asarray(initial, dtype=np.float64)
    d = best[:, 5] - initial[:, 5]
    back = (d < 0) & (np.abs(d) > drop * initial[:, 5])
    best[back] = initial[back]
    return best


def BH_rawAlignmentsApply(INPUT_GEOMETRY: dict, BEST_ANGLES: dict, SAMPLING, nPeaks: int, rotConvention="Bah",
                          updateWeights=False, updateClassByBestReferenceScore=False) -> dict:
    """metaData/BH_rawAlignmentsApply.m:40-89 over every tomogram's 26*nPeaks
    column geometry: particles not ignored (class column 26 != -9999) take
    the row of their id (column 4); without one they are ignored; else
    rotation (17:25, column-major) <- old * BH_defineMatrix(angles, conv,
    'inv'), position (11:13) += shift * SAMPLING, column 1 <- CCC, column 2
    <- column 7 of the row (updateWeights), column 26 <- the row's class
    (updateClassByBestReferenceScore)."""
    out = {}
    for tomo, geom in INPUT_GEOMETRY.items():
        pos = np.array(geom, dtype=np.float64, copy=True)
        new = np.asarray(BEST_ANGLES.get(tomo, np.zeros((0, 10 * nPeaks))), dtype=np.float64)
        if new.size:
            new = new.reshape(new.shape[0], -1)
            for p in range(nPeaks):
                o, q = 26 * p, 10 * p
                live = np.flatnonzero(pos[:, 25 + o] != -9999)
                ids = new[:, 1]
                order = np.argsort(ids, kind="stable")
                k = np.searchsorted(ids[order], pos[live, 3 + o])
                k = np.minimum(k, len(order) - 1)
                found = ids[order][k] == pos[live, 3 + o]
                pos[live[~found], 25 + o] = -9999
                idx, rows = live[found], new[order[k[found]]]
                if idx.size == 0:
                    continue
                cls = np.stack([BH_defineMatrix(a, rotConvention, "inv") for a in rows[:, 2 + q:5 + q]])
                old = pos[idx, 16 + o:25 + o].reshape(-1, 3, 3).transpose(0, 2, 1)   # column-major
                pos[idx, 16 + o:25 + o] = (old @ cls).transpose(0, 2, 1).reshape(-1, 9)
                pos[idx, 10 + o:13 + o] += rows[:, 7 + q:10 + q] * SAMPLING
                pos[idx, 0 + o] = rows[:, 5 + q]
                if updateWeights:
                    pos[idx, 1 + o] = rows[:, 6 + q]
                if updateClassByBestReferenceScore:
                    pos[idx, 25 + o] = rows[:, 0 + q]
        out[tomo] = pos
    return out


def align_tomogram(tomo: ResidentTomogram, positionList, ref, bandpass, grid_pts, sampling=1, bin_shift=(0, 0, 0),
                   mask_radius=None, wedges=None, class_id=1, table="c64", refine_iters=12, peak_radius=6,
                   aligner: BatchAligner | None = None):
    """One tomogram's alignment pass on the batched path: every particle not
    ignored is cut from the resident tomogram at positionList(:, 11:13) /
    sampling + bin_shift (one emc_extract_subtomos), searched over grid_pts
    (N x (phi, theta, psi)) around its orientation positionList(:, 17:25),
    refined on the device, and reported as the reference's rows (one per
    particle, in geometry order; 'noUse' windows give none). ``wedges``: None
    or a per-particle list of (n, n, n) weights. Returns (rows (m, 10),
    geometry row index of each row)."""
    pos = np.asarray(positionList, dtype=np.float64)
    n = int(np.asarray(ref).shape[0]) if aligner is None else aligner.n
    mask_radius = mask_radius if mask_radius is not None else 0.25 * n
    live = np.flatnonzero(pos[:, 25] != -9999)
    centres = pos[live, 10:13] / sampling + np.asarray(bin_shift, dtype=np.float64)
    win, valid, shift_val = extract_subtomos(tomo, centres, (n, n, n), mask_radius)
    use = np.flatnonzero(valid)
    if use.size == 0:
        return np.zeros((0, 10)), np.zeros(0, np.int64)
    vox = n * n * n * 4
    parts = [win.view((n, n, n), np.float32, int(i) * vox) for i in use]
    starts = [pos[live[i], 16:25].reshape(3, 3, order="F") for i in use]
    w = [None] * use.size if wedges is None else [wedges[live[i]] for i in use]
    al = aligner or BatchAligner(n, ref, bandpass, peak_radius=peak_radius, table=table, ctx=tomo.raw.ctx)
    al.set_particles(parts, w, starts)
    pts = np.asarray(grid_pts, dtype=np.float64).reshape(-1, 3)
    score, shift = al.search(pts)
    P = use.size
    best = np.argmax(score, axis=1)
    from .workload import local_matrix
    R0 = np.stack([al.starts[i] @ local_matrix(*pts[best[i]]) for i in range(P)])
    s0 = shift[np.arange(P), best].astype(np.float64)
    if refine_iters > 0:
        R, s, cc = al.refine(R0, s0, iters=refine_iters)
    else:
        R, s, cc = R0, s0, score[np.arange(P), best].astype(np.float64)
    rel = np.einsum("pji,pjk->pik", al.starts, R)            # start^T R
    sv = shift_val[use]
    rows = np.zeros((P, 10))
    rows[:, 0] = class_id
    rows[:, 1] = pos[live[use], 3]
    rows[:, 2:5] = bah_angles_from_matrix(rel)
    rows[:, 5] = cc
    rows[:, 6] = 1.0
    rows[:, 7:10] = s - sv
    ident = np.flatnonzero(np.all(np.abs(pts) < 1e-12, axis=1))
    if ident.size:
        i0 = ident[0]
        init = np.zeros_like(rows)
        init[:, 0:2] = rows[:, 0:2]
        init[:, 5] = score[:, i0]
        init[:, 6] = 1.0
        init[:, 7:10] = shift[:, i0] - sv
        rows = revert_drops(rows, init)
    return rows, live[use]


def angle_error_deg(Ra, Rb) -> float:
    c = (np.trace(np.asarray(Ra).T @ np.asarray(Rb)) - 1.0) / 2.0
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))
