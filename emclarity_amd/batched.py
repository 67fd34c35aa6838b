"""Batched subtomogram alignment with adjoint (gradient) refinement.

Host driver of emc_align_batched / emc_align_eval_grad (see
include/emc_hip.h): particles of one box size are prepared once (spectrum x
bandpass x their own wedge), scored against a block of orientations in one
launch, and the best grid pose of every particle is then refined by
quasi-Newton ascent on the normalised CC using the hand-written adjoint.
The orientation search replaces the per-orientation loop of
alignment/BH_alignRaw3d_v2.m:1119-1419; the gradient refinement replaces its
dense local grid (:1448-1644).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .device import DeviceArray, as_device, default_context

_G = [np.array([[0.0, 0, 0], [0, 0, -1], [0, 1, 0]]),
      np.array([[0.0, 0, 1], [0, 0, 0], [-1, 0, 0]]),
      np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 0]])]


def rodrigues(w):
    """exp([w]x) for rotation vectors w (..., 3)."""
    w = np.asarray(w, dtype=np.float64)
    th = np.linalg.norm(w, axis=-1, keepdims=True)
    k = np.where(th > 1e-15, w / np.where(th > 1e-15, th, 1.0), 0.0)
    K = np.zeros(w.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2] = -k[..., 2], k[..., 1]
    K[..., 1, 0], K[..., 1, 2] = k[..., 2], -k[..., 0]
    K[..., 2, 0], K[..., 2, 1] = -k[..., 1], k[..., 0]
    s = np.sin(th)[..., None]
    c = (1.0 - np.cos(th))[..., None]
    return np.eye(3) + s * K + c * (K @ K)


def euler_derivatives(A, phi, theta, psi):
    """R = A Rz(phi) Rx(theta) Rz(psi - phi) (BH_defineMatrix 'inv' of
    [phi, theta, psi - phi]) and dR/dphi, dR/dtheta, dR/dpsi (per radian)."""
    def rz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])

    def drz(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[-s, -c, 0], [c, -s, 0], [0, 0, 0.0]])

    def rx(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])

    def drx(a):
        c, s = math.cos(a), math.sin(a)
        return np.array([[0.0, 0, 0], [0, -s, -c], [0, c, -s]])

    p, t, q = (math.radians(v) for v in (phi, theta, psi))
    R = A @ rz(p) @ rx(t) @ rz(q - p)
    dphi = A @ (drz(p) @ rx(t) @ rz(q - p) - rz(p) @ rx(t) @ drz(q - p))
    dtheta = A @ rz(p) @ drx(t) @ rz(q - p)
    dpsi = A @ rz(p) @ rx(t) @ drz(q - p)
    return R, [dphi, dtheta, dpsi]


def support_radius_of(bandpass, n, thresh=1e-4):
    """Largest frequency radius (index units) where the bandpass (FFT order)
    exceeds `thresh`, capped below n/2."""
    bp = np.asarray(bandpass)
    f2 = (np.fft.fftfreq(n) * n).astype(np.int32) ** 2
    r2 = f2[:, None, None] + f2[None, :, None] + f2[None, None, :]
    m = bp > thresh
    rs = math.sqrt(float(r2[m].max())) if m.any() else n / 4
    return min(rs + 0.5, n / 2 - 1.01)


TABLE_FORMATS = {"c64": 0, "f16": 1}


class BatchAligner:
    """Particles of one box size n against one reference. ``table`` is the
    reference table format: "c64" (complex64) or "f16" (half-precision
    parts, half the gather bytes)."""

    def __init__(self, n, ref, bandpass, support_radius=None, peak_radius=6, ctx=None, table="c64", tuning=None):
        self.ctx = ctx or default_context()
        self.n = int(n)
        # launch shapes from the autotuner's table (tuning None: the shipped
        # table; {}: the library defaults)
        from .autotune import params_for
        self._tune = params_for(self.n, table, tuning)
        self.peak_radius = int(peak_radius)
        if table not in TABLE_FORMATS:
            raise ValueError(f"table must be one of {sorted(TABLE_FORMATS)}")
        self.table = table
        self._fmt = TABLE_FORMATS[table]
        if support_radius:
            self.support = float(support_radius)
        else:
            bp_host = bandpass.get() if isinstance(bandpass, DeviceArray) else np.asarray(bandpass)
            self.support = support_radius_of(bp_host, n)
        self.bandpass = as_device(bandpass, np.float32, self.ctx)
        n2 = 2 * self.n
        texel = np.complex64 if table == "c64" else np.float32   # f16: two halves per texel
        self.F2 = DeviceArray((n2 + 1, n2 + 1, self.n + 2, 4), texel, self.ctx)
        r = as_device(ref, np.float32, self.ctx)
        _lib.call("emc_batch_reference_spectrum", self.ctx.handle, self.n, r.vp, self._fmt, self.F2.vp)
        self.n_part = 0

    @property
    def nh(self):
        return (self.n // 2 + 1) * self.n * self.n

    def set_particles(self, particles, wedges, starts, mask=None):
        """particles / wedges: lists of (n,n,n) arrays (wedge None = 1);
        starts: list of 3x3 start matrices; mask: an (n,n,n) centred
        real-space mask applied to every particle first (the reference's
        volMask; None = none)."""
        P = len(particles)
        self.n_part = P
        self.Q = DeviceArray((P, self.nh), np.complex64, self.ctx)
        self.N = DeviceArray((P, self.nh), np.float32, self.ctx)
        self.A = DeviceArray((P,), np.float64, self.ctx)
        # adjoint inputs (36 matrix + 3 shift floats per particle) and outputs,
        # reused by every eval_grad call on this particle set
        self._h_in = np.zeros(39 * P, np.float32)   # [P x 36 matrices | P x 3 shifts]
        self._d_in = DeviceArray((39 * P,), np.float32, self.ctx)
        self._d_out = DeviceArray((P, 7), np.float64, self.ctx)
        self.starts = np.stack([np.asarray(s, dtype=np.float64).reshape(3, 3) for s in starts])
        self.d_starts = DeviceArray.from_numpy(
            np.ascontiguousarray(self.starts.transpose(0, 2, 1).reshape(P, 9)), ctx=self.ctx, wait=False)
        # device copies of host inputs stay referenced until the call returns
        pd = [as_device(p, np.float32, self.ctx) for p in particles]
        wd = [None if w is None else as_device(w, np.float32, self.ctx) for w in wedges]
        hp = (C.c_void_p * P)(*[a.ptr for a in pd])
        hw = (C.c_void_p * P)(*[None if a is None else a.ptr for a in wd])
        if mask is None:
            _lib.call("emc_batch_prepare_particles", self.ctx.handle, self.n, P, hp, self.bandpass.vp, hw,
                      self.Q.vp, self.N.vp, self.A.vp)
        else:
            md = as_device(mask, np.float32, self.ctx)
            if md.shape != (self.n,) * 3:
                raise ValueError(f"mask must be ({self.n},)*3")
            _lib.call("emc_batch_prepare_particles_masked", self.ctx.handle, self.n, P, hp, md.vp,
                      self.bandpass.vp, hw, self.Q.vp, self.N.vp, self.A.vp)

    # ------------------------------------------------------------ search
    def _apply_tuning(self):
        if self._tune:
            from .autotune import apply
            apply(self.ctx, self.n, self._tune)

    def search_device(self, angles_pts, support_radius=None):
        """emc_align_batched, results left on the device: score (P, O) and
        shift (P, O, 3) DeviceArrays (stream-ordered, no wait).
        support_radius: the frequencies scored (default: the bandpass
        support)."""
        self._apply_tuning()
        r = self.support if support_radius is None else float(support_radius)
        pts = np.ascontiguousarray(np.asarray(angles_pts, dtype=np.float64).reshape(-1, 3))
        O = pts.shape[0]
        score = DeviceArray((self.n_part, O), np.float32, self.ctx)
        shift = DeviceArray((self.n_part, O, 3), np.float32, self.ctx)
        _lib.call("emc_align_batched", self.ctx.handle, self.n, self.n_part, self._fmt, self.F2.vp, self.Q.vp,
                  self.N.vp, self.A.vp, self.d_starts.vp, pts.ctypes.data_as(C.POINTER(C.c_double)), O,
                  self.peak_radius, r, score.vp, shift.vp)
        return score, shift

    def search(self, angles_pts, support_radius=None):
        """Scores (P, O) and shifts (P, O, 3) of every (particle, orientation)."""
        score, shift = self.search_device(angles_pts, support_radius)
        return score.get(), shift.get()

    def pick(self, score, shift, d_pts):
        """Best orientation per particle on the device (emc_align_pick):
        (R0 (P, 9) row-major FP64, s0 (P, 3), best index (P,), best score (P,))."""
        P, O = score.shape
        R0 = DeviceArray((P, 9), np.float64, self.ctx)
        s0 = DeviceArray((P, 3), np.float64, self.ctx)
        best = DeviceArray((P,), np.int32, self.ctx)
        bs = DeviceArray((P,), np.float32, self.ctx)
        _lib.call("emc_align_pick", self.ctx.handle, P, O, score.vp, shift.vp, self.d_starts.vp, d_pts.vp, R0.vp,
                  s0.vp, best.vp, bs.vp)
        return R0, s0, best, bs

    # ------------------------------------------------------------ adjoint
    def eval_grad(self, R, dR, s):
        """CC and gradient at per-particle poses: R (P,3,3), dR (P,3,3,3)
        derivatives w.r.t. three rotation parameters, s (P,3)."""
        self._apply_tuning()
        P = self.n_part
        h = self._h_in
        hm = h[:36 * P].reshape(P, 36)
        hm[:, :9] = 2.0 * np.transpose(R, (0, 2, 1)).reshape(P, 9)
        hm[:, 9:] = 2.0 * np.transpose(dR, (0, 1, 3, 2)).reshape(P, 27)
        h[36 * P:] = np.asarray(s).reshape(3 * P)
        _lib.call("emc_memcpy_h2d", self.ctx.handle, self._d_in.vp, h.ctypes.data_as(C.c_void_p),
                  C.c_size_t(h.nbytes))
        d_m = self._d_in.ptr
        d_s = self._d_in.ptr + 36 * 4 * P
        _lib.call("emc_align_eval_grad", self.ctx.handle, self.n, P, self._fmt, self.F2.vp, self.Q.vp, self.N.vp,
                  self.A.vp, C.c_void_p(d_m), C.c_void_p(d_s), self.support, self._d_out.vp)
        o = self._d_out.get()
        return o[:, 0], o[:, 1:]

    def _eval_local(self, R, s):
        """Gradient w.r.t. right-multiplied small rotations R exp([w]x)."""
        dR = np.stack([R @ g for g in _G], axis=1)
        return self.eval_grad(R, dR, s)

    def refine_device(self, R, s, iters=12, step_angle=0.02, step_shift=0.3, support_radius=None):
        """emc_align_refine on device arrays R (P, 9) row-major and s (P, 3),
        FP64, updated in place; returns the final CC as a device array.
        support_radius: the frequencies the CC is evaluated over (default:
        the bandpass support)."""
        self._apply_tuning()
        cc = DeviceArray((self.n_part,), np.float64, self.ctx)
        r = self.support if support_radius is None else float(support_radius)
        _lib.call("emc_align_refine", self.ctx.handle, self.n, self.n_part, self._fmt, self.F2.vp, self.Q.vp,
                  self.N.vp, self.A.vp, R.vp, s.vp, int(iters), float(step_angle), float(step_shift), r, cc.vp)
        return cc

    def refine(self, R0, s0, iters=12, step_angle=0.02, step_shift=0.3):
        """Batched quasi-Newton ascent of the CC over (rotation increment,
        shift) for every particle, on the device (emc_align_refine: iters + 1
        adjoint evaluations, one result copy). Returns (R, s, cc)."""
        P = self.n_part
        R = DeviceArray.from_numpy(np.ascontiguousarray(np.asarray(R0, dtype=np.float64).reshape(P, 9)),
                                   ctx=self.ctx)
        s = DeviceArray.from_numpy(np.ascontiguousarray(np.asarray(s0, dtype=np.float64).reshape(P, 3)),
                                   ctx=self.ctx)
        cc = self.refine_device(R, s, iters, step_angle, step_shift)
        return R.get().reshape(P, 3, 3), s.get(), cc.get()


def support_voxels(n, support):
    """Half-grid frequencies inside the support sphere (what the fused
    kernel evaluates per unit)."""
    f = np.fft.fftfreq(n) * n
    kz, ky = np.meshgrid(f, f, indexing="ij")
    kx = np.arange(n // 2 + 1)
    r2 = kz[:, :, None] ** 2 + ky[:, :, None] ** 2 + kx[None, None, :] ** 2
    return int(np.count_nonzero(r2 <= support * support))


def fused_schedule(n, support, tile=16, width=256):
    """Column schedule of the fused search kernel (restated from its host
    builder): columns (kx, ky, h) of the half disk kx^2 + ky^2 <= r^2, tiled
    in tile x tile (kx, ky) patches, consecutive patches merged up to `width`
    columns per pass. Returns (passes, pieces): each pass is a list of
    columns; pieces counts the runs of one kx inside the passes."""
    k2 = np.float32(support * support)
    tiles = {}
    for kx in range(0, n // 2 + 1):
        if np.float32(kx * kx) > k2:
            break
        for ky in range(-(n // 2 - 1), n // 2):
            if np.float32(kx * kx + ky * ky) > k2:
                continue
            h = 0
            while h + 1 <= n // 2 - 1 and np.float32(kx * kx + ky * ky + (h + 1) ** 2) <= k2:
                h += 1
            tiles.setdefault((kx // tile, (ky + 128) // tile), []).append((kx, ky, h))
    passes, cur = [], []
    for key in sorted(tiles):
        if len(cur) + len(tiles[key]) > width:
            passes.append(cur)
            cur = []
        cur += tiles[key]
    if cur:
        passes.append(cur)
    pieces = sum(len({c[0] for c in g}) for g in passes)
    return passes, pieces


def fused_flops_per_unit(n, support, peak_radius):
    """FP32 operations of one (particle, orientation) in align_fused_kernel,
    counted from its arithmetic (an FMA counts 2):
      per support frequency: j = M k from the column base (3 FMA), trilinear
      (7 complex lerps, 28), X = Q conj(f) (6), norm term (4), z-pruned DFT
      (D complex MACs, 8 D);
      per column: y-pruned DFT (D^2 complex MACs, 8 D^2);
      per plane piece: Hermitian x expansion (D^3 real MACs of a complex
      product, 4 D^3)."""
    D = 2 * peak_radius + 1
    passes, pieces = fused_schedule(n, support)
    cols = sum(len(g) for g in passes)
    v = sum(2 * h + 1 for g in passes for (_, _, h) in g)
    return v * (6 + 28 + 6 + 4 + 8 * D) + cols * 8 * D * D + pieces * 4 * D ** 3


def fused_samples_per_unit(n, support):
    """Trilinear reference samples one (particle, orientation) takes."""
    passes, _ = fused_schedule(n, support)
    return sum(2 * h + 1 for g in passes for (_, _, h) in g)


def search_stage(aligner: BatchAligner, grid_pts, after=None, support_frac=1.0):
    """Queue the coarse search over grid_pts and the pick of each particle's
    best orientation (first maximum) on the aligner's stream; no host round
    trip. ``after``: a Context whose queued work the search waits for, so
    searches of several aligners on their own streams run one after another
    while the rest of their work overlaps. ``support_frac``: score the grid
    over that fraction of the support radius (a coarse grid only has to find
    the refinement's basin; the refinement uses the full support)."""
    pts = np.ascontiguousarray(np.asarray(grid_pts, dtype=np.float64).reshape(-1, 3))
    d_pts = DeviceArray.from_numpy(pts, ctx=aligner.ctx, wait=False)
    if after is not None:
        aligner.ctx.wait_for(after)
    score, shift = aligner.search_device(pts, support_radius=support_frac * aligner.support)
    R, s, best, bs = aligner.pick(score, shift, d_pts)
    return {"aligner": aligner, "P": aligner.n_part, "R": R, "s": s, "cc": None, "best": best, "best_score": bs,
            "keep": (score, shift, d_pts)}


def refine_stage(h, refine_iters=12, coarse=None):
    """Queue the adjoint refinement from a search_stage handle's picks.
    coarse = (fraction, iterations): first that many iterations with the CC
    evaluated inside fraction x the support radius (fewer frequencies, a
    smoother score), then the rest over the full support, each stage from a
    fresh quasi-Newton state at the previous stage's pose."""
    al = h["aligner"]
    n_full = refine_iters
    if coarse is not None and coarse[1] > 0 and refine_iters > 0:
        frac, k = float(coarse[0]), min(int(coarse[1]), refine_iters)
        h["cc"] = al.refine_device(h["R"], h["s"], iters=k, support_radius=frac * al.support)
        n_full = refine_iters - k
    if n_full > 0:
        h["cc"] = al.refine_device(h["R"], h["s"], iters=n_full)
    return h


def align_batch_async(aligner: BatchAligner, grid_pts, refine_iters=12):
    """search_stage + refine_stage; collect() waits for the results."""
    return refine_stage(search_stage(aligner, grid_pts), refine_iters)


def collect(h):
    """(R (P,3,3), shift (P,3), cc (P,), grid best index (P,)) of an
    align_batch_async handle."""
    P = h["P"]
    cc = h["cc"].get() if h["cc"] is not None else h["best_score"].get().astype(np.float64)
    return h["R"].get().reshape(P, 3, 3), h["s"].get(), cc, h["best"].get().astype(np.int64)


def align_batch(aligner: BatchAligner, grid_pts, refine_iters=12):
    """Coarse search over grid_pts, best per particle (first maximum), then
    adjoint refinement. Returns (R (P,3,3), shift (P,3), cc (P,), grid best
    index (P,))."""
    return collect(align_batch_async(aligner, grid_pts, refine_iters))


# ---------------------------------------------------------------- checks


def euler_eval(aligner: BatchAligner, local_pts, shifts):
    """CC and gradient of every particle at pose start_i R(phi, theta, psi)
    (local Euler angles in degrees, BH_defineMatrix 'inv' of [phi, theta,
    psi - phi]) and shift s_i: (cc (P,), grad (P, 6) = d/dphi, d/dtheta,
    d/dpsi per radian, then d/dsx, d/dsy, d/dsz per pixel)."""
    P = aligner.n_part
    pts = np.asarray(local_pts, dtype=np.float64).reshape(P, 3)
    R = np.empty((P, 3, 3))
    dR = np.empty((P, 3, 3, 3))
    for i in range(P):
        R[i], d = euler_derivatives(aligner.starts[i], *pts[i])
        dR[i] = np.stack(d)
    return aligner.eval_grad(R, dR, np.asarray(shifts, dtype=np.float64).reshape(P, 3))


def finite_difference_check(aligner: BatchAligner, local_pts, shifts, h=1e-3, hs=1e-2):
    """Central differences of the device CC in each of the six parameters
    (h radians for the Euler angles, hs pixels for the shifts), every
    particle at once: 12 batched evaluations. Returns (grad (P, 6) from the
    adjoint, fd (P, 6))."""
    P = aligner.n_part
    pts = np.asarray(local_pts, dtype=np.float64).reshape(P, 3)
    s = np.asarray(shifts, dtype=np.float64).reshape(P, 3)
    _, g = euler_eval(aligner, pts, s)
    fd = np.empty((P, 6))
    for i in range(3):
        e = np.zeros(3)
        e[i] = math.degrees(h)
        cp, _ = euler_eval(aligner, pts + e, s)
        cm, _ = euler_eval(aligner, pts - e, s)
        fd[:, i] = (cp - cm) / (2 * h)
    for i in range(3):
        e = np.zeros(3)
        e[i] = hs
        cp, _ = euler_eval(aligner, pts, s + e)
        cm, _ = euler_eval(aligner, pts, s - e)
        fd[:, 3 + i] = (cp - cm) / (2 * hs)
    return g, fd


def judge(aligner: BatchAligner, R_grid, s_grid, R, s, cc):
    """The plain grid as fallback and judge of the gradient refinement: the
    CC of each particle's grid pick is evaluated over the full support (the
    refinement's own score) and a particle whose refined CC is lower keeps
    its grid pose. Returns (R, s, cc, fell_back (P,) bool)."""
    P = aligner.n_part
    Rg = np.asarray(R_grid, dtype=np.float64).reshape(P, 3, 3)
    sg = np.asarray(s_grid, dtype=np.float64).reshape(P, 3)
    cg, _ = aligner._eval_local(Rg, sg)
    back = cg > np.asarray(cc)
    R = np.where(back[:, None, None], Rg, np.asarray(R).reshape(P, 3, 3))
    s = np.where(back[:, None], sg, np.asarray(s).reshape(P, 3))
    return R, s, np.where(back, cg, cc), back
