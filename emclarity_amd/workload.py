"""Seeded alignment workloads for the benchmark and the accuracy tests.

Batched alignment workload (mixed-size subtomograms): particles of box 64^3 /
96^3 / 128^3 in 50/30/20 proportion at 2.5 A/px, each the reference rotated
and shifted (U[-3,3] px), with its own missing wedge (max tilt U[50,70] deg)
and defocus (U[1.5,4.0] um), SNR 0.1. Each particle carries a start
orientation A such that the true pose lies inside the local search
[theta_max 15, step 3, psi_max 15, step 3] (1,045 orientations) around A:
R_true = A * BH_defineMatrix([phi, theta, psi - phi], 'Bah', 'inv') for an
off-grid (phi, theta, psi) with theta <= 12 deg and |psi| <= 12 deg.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import synthetic as S
from .geometry import BH_defineMatrix, eulerSearch


@dataclass
class Particle:
    n: int
    data: np.ndarray          # (n,n,n) float32 subtomogram
    start: np.ndarray         # 3x3 start orientation A
    truth_R: np.ndarray       # 3x3 true RotMat
    truth_local: tuple        # (phi, theta, psi) of the truth inside the local search
    shift: np.ndarray         # true shift (x,y,z)
    max_tilt: float
    defocus_um: float
    tlt: np.ndarray = field(repr=False, default=None)
    signal: np.ndarray = field(repr=False, default=None)   # noise-free wedge x CTF volume (kept on request)
    noise_sd: float = 0.0


@dataclass
class SizeClass:
    n: int
    ref: np.ndarray           # reference (n,n,n)
    bandpass: np.ndarray      # FFT order (n,n,n)
    peak_mask: np.ndarray     # centred sphere
    vol_mask: np.ndarray


def random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def local_matrix(phi, theta, psi):
    return BH_defineMatrix([phi, theta, psi - phi], "Bah", "inv")


def make_size_class(n, seed=100, pixel_A=2.5):
    rng = np.random.default_rng(seed + n)
    blobs = S.make_blobs(rng, 40, radius=0.3 * n, sigma=(2.0 * n / 64, 5.0 * n / 64))
    from .masking import BH_bandpass3d
    bp = BH_bandpass3d((n, n, n), 1e-3, 40.0, 8.0, "GPU", pixel_A).get()
    return SizeClass(n, S.render(blobs, n).astype(np.float32), bp,
                     S.sphere_mask(n, n // 8 + 4), S.sphere_mask(n, 0.45 * n)), blobs


def _local_truth(rng):
    """Off-grid local truth inside the +-15 deg search: (phi, theta, psi)."""
    theta = rng.uniform(0.5, 12.0)
    phi = rng.uniform(0.0, 360.0)
    psi = rng.uniform(-12.0, 12.0)
    return phi, theta, psi


def make_particle(rng, sc: SizeClass, blobs, pixel_A=2.5, snr=0.1, keep_signal=False):
    n = sc.n
    A = random_rotation(rng)
    phi, theta, psi = _local_truth(rng)
    Rt = A @ local_matrix(phi, theta, psi)
    shift = rng.uniform(-3.0, 3.0, 3)
    max_tilt = rng.uniform(50.0, 70.0)
    df = rng.uniform(1.5, 4.0)
    clean = S.render(blobs, n, Rt, shift)
    sig = S.degrade(clean, rng, max_tilt, pixel_A, df, None, ctf=True)
    sd = math.sqrt(sig.var() / snr)
    data = (sig + rng.standard_normal(sig.shape) * sd).astype(np.float32)
    T = S.tilt_table(max_tilt, 3.0, df, pixel_A)
    return Particle(n, data, A, Rt, (phi, theta, psi), shift, max_tilt, df, T,
                    sig.astype(np.float32) if keep_signal else None, sd)


def copy_particle(rng, p: Particle):
    """Another subtomogram of the same object, pose, shift, wedge and
    defocus as ``p`` (made with keep_signal=True): a fresh noise realisation
    at the same SNR and a new start orientation, the truth again an off-grid
    point of the +-15 deg local search around it."""
    if p.signal is None:
        raise ValueError("the particle was made without keep_signal")
    phi, theta, psi = _local_truth(rng)
    A = p.truth_R @ local_matrix(phi, theta, psi).T
    data = p.signal + rng.standard_normal(p.signal.shape, dtype=np.float32) * np.float32(p.noise_sd)
    return Particle(p.n, data, A, p.truth_R, (phi, theta, psi), p.shift, p.max_tilt, p.defocus_um, p.tlt)


# ------------------------------------------------------------------
# Fast generation of many distinct particles (the 2,000-particle workload:
# every particle its own object pose, shift, wedge and defocus) and the
# tomograms they are cut from.


def render_fast(blobs: S.Blobs, n: int, R=None, shift=(0.0, 0.0, 0.0)) -> np.ndarray:
    """synthetic.render as one contraction: each blob is the outer product of
    three 1-D Gaussians, so the volume is sum_b gz_b (x) gy_b (x) gx_b."""
    R = np.eye(3) if R is None else np.asarray(R, dtype=np.float64)
    cen = blobs.centres @ R.T + np.asarray(shift, dtype=np.float64)
    ax = np.arange(n, dtype=np.float64) - n // 2
    s2 = 2.0 * blobs.sigma[:, None] ** 2
    gx = np.exp(-((ax[None, :] - cen[:, 0:1]) ** 2) / s2)
    gy = np.exp(-((ax[None, :] - cen[:, 1:2]) ** 2) / s2)
    gz = np.exp(-((ax[None, :] - cen[:, 2:3]) ** 2) / s2) * blobs.amp[:, None]
    return np.einsum("bz,by,bx->zyx", gz, gy, gx, optimize=True)


class Degrader:
    """synthetic.degrade on the half grid in FP32 for one box size: the
    missing wedge about Y (|kz| <= tan(max tilt) |kx|) times |CTF| is even in
    k, so the real part of the full-grid product equals the half-grid
    inverse transform."""

    def __init__(self, n: int, pixel_A: float = 2.5, kv=300.0, cs_mm=2.7, amp=0.1):
        import scipy.fft as sfft
        self._fft = sfft
        self.n = n
        f = np.fft.fftfreq(n).astype(np.float64)
        fx = np.fft.rfftfreq(n).astype(np.float64)
        kz, ky, kx = np.meshgrid(f, f, fx, indexing="ij")
        self.akx = np.abs(kx)
        self.akz = np.abs(kz)
        # the CTF depends on |k| only: evaluate it once per distinct integer
        # radius^2 and gather
        iz, iy, ix = np.meshgrid(np.fft.fftfreq(n) * n, np.fft.fftfreq(n) * n, np.arange(n // 2 + 1), indexing="ij")
        r2 = np.rint(ix * ix + iy * iy + iz * iz).astype(np.int64)
        self.r2u, self.r2i = np.unique(r2, return_inverse=True)
        self.r2i = self.r2i.reshape(r2.shape).astype(np.int32)
        lam = 12.2643247 / math.sqrt(kv * 1e3 * (1.0 + kv * 1e3 * 0.978466e-6))
        k2 = self.r2u / (n * n * pixel_A * pixel_A)
        self.c4 = math.pi * 0.5 * cs_mm * 1e7 * lam ** 3 * k2 * k2
        self.c2 = math.pi * lam * k2
        self.c0 = math.atan(amp / math.sqrt(1 - amp * amp))

    def filter(self, max_tilt, defocus_um):
        chi = self.c4 - self.c2 * (defocus_um * 1e4) - self.c0
        ctf = np.abs(np.sin(chi)).astype(np.float32)[self.r2i]
        w = (self.akz <= math.tan(math.radians(max_tilt)) * self.akx + 1e-12)
        return np.where(w, ctf, np.float32(0.0))

    def __call__(self, vol, rng, max_tilt, defocus_um, snr):
        F = self._fft.rfftn(np.asarray(vol, dtype=np.float32), workers=1)
        sig = self._fft.irfftn(F * self.filter(max_tilt, defocus_um), s=vol.shape, workers=1)
        sd = math.sqrt(float(sig.var()) / snr)
        return (sig + rng.standard_normal(sig.shape, dtype=np.float32) * np.float32(sd)).astype(np.float32)


def make_particle_fast(rng, sc: SizeClass, blobs, degrader: Degrader, pixel_A=2.5, snr=0.1, max_tilt=None):
    """make_particle with render_fast and the half-grid degradation; the same
    distributions (start, local truth, shift, max tilt, defocus). max_tilt:
    a fixed wedge (degrees) instead of U[50, 70]."""
    n = sc.n
    A = random_rotation(rng)
    phi, theta, psi = _local_truth(rng)
    Rt = A @ local_matrix(phi, theta, psi)
    shift = rng.uniform(-3.0, 3.0, 3)
    drawn = rng.uniform(50.0, 70.0)
    max_tilt = drawn if max_tilt is None else float(max_tilt)
    df = rng.uniform(1.5, 4.0)
    data = degrader(render_fast(blobs, n, Rt, shift), rng, max_tilt, df, snr)
    T = S.tilt_table(max_tilt, 3.0, df, pixel_A)
    return Particle(n, data, A, Rt, (phi, theta, psi), shift, max_tilt, df, T)


@dataclass
class CellTomogram:
    """Particles of one box size n laid in an int16 tomogram, particle i in
    cell (i % gx, (i // gx) % gy, i // (gx gy)) of n^3 voxels. ``centres``
    are the 1-based centres (x, y, z) BH_isWindowValid cuts the n^3 windows
    around (the cells); ``tlt`` the particles' tilt tables with the defocus
    raised by each particle's height in the tomogram, so that BH_weightMaskMex
    (which lowers it by that height, masking/BH_weightMaskMex.m:33-37) gives
    the particle's own defocus at zero tilt; ``recon`` the recon geometry
    rows [[nx, ny, nz], [0, 0, 0]]."""
    n: int
    data: np.ndarray
    scale: float
    centres: np.ndarray
    tlt: list
    recon: np.ndarray


def cell_tomogram(particles, n, pixel_A=2.5) -> CellTomogram:
    P = len(particles)
    gx = max(1, int(math.ceil(P ** (1.0 / 3.0))))
    gy = gx
    gz = int(math.ceil(P / (gx * gy)))
    dims = (gz * n, gy * n, gx * n)
    peak = max(float(np.max(np.abs(p.data))) for p in particles)
    scale = 30000.0 / peak
    vol = np.zeros(dims, np.int16)
    cen = np.zeros((P, 3))
    tlts = []
    oz = math.ceil((dims[0] + 1) / 2.0)
    for i, p in enumerate(particles):
        cx, cy, cz = i % gx, (i // gx) % gy, i // (gx * gy)
        vol[cz * n:(cz + 1) * n, cy * n:(cy + 1) * n, cx * n:(cx + 1) * n] = np.rint(p.data * scale).astype(np.int16)
        # window index n//2 of the cell is the 1-based centre floor(c)
        cen[i] = (cx * n + n // 2 + 1, cy * n + n // 2 + 1, cz * n + n // 2 + 1)
        T = p.tlt.copy()
        T[:, 14] -= (cen[i, 2] - oz) * pixel_A * 1e-10       # column 15 holds -defocus (m)
        tlts.append(T)
    recon = np.array([[dims[2], dims[1], dims[0]], [0.0, 0.0, 0.0]])
    return CellTomogram(n, vol, scale, cen, tlts, recon)


def mixed_sizes(count, fractions=((64, 0.5), (96, 0.3), (128, 0.2))):
    sizes = []
    for n, f in fractions:
        sizes += [n] * int(round(count * f))
    while len(sizes) < count:
        sizes.append(fractions[0][0])
    return sizes[:count]


def angle_error_deg(Ra, Rb):
    c = (np.trace(Ra.T @ Rb) - 1.0) / 2.0
    return math.degrees(math.acos(max(-1.0, min(1.0, c))))


def accuracy(parts, Rs, shifts):
    """Alignment accuracy against the ground truth: how many particles land
    within 1 degree (orientation) and 0.5 px (largest shift component), and
    the median errors."""
    errs = np.array([(angle_error_deg(R, p.truth_R), float(np.max(np.abs(np.asarray(s) - p.shift))))
                     for p, R, s in zip(parts, Rs, shifts)])
    ok = int(np.sum((errs[:, 0] <= 1.0) & (errs[:, 1] <= 0.5)))
    return {"within_1deg_0p5px": ok, "of": len(parts), "fraction": round(ok / len(parts), 4),
            "median_angle_err_deg": round(float(np.median(errs[:, 0])), 4),
            "median_shift_err_px": round(float(np.median(errs[:, 1])), 4)}


def local_grid():
    """The 1,045-orientation local search ([15,3,15,3], eulerSearch)."""
    return eulerSearch("C1", 15, 3, 15, 3)


def gradient_batch(count=256, n=128, seed=31, workers=8, max_tilt=60.0):
    """The differentiable-alignment workload: ``count`` particles of n^3 at
    SNR 0.1, a +-``max_tilt`` wedge about Y, defocus U[1.5, 4.0] um each,
    shifts U[-3, 3] px, the truth off-grid inside the +-15 deg local search
    around each start. Seeded per particle. Returns (size class, particles)."""
    import concurrent.futures as cf
    sc, blobs = make_size_class(n)
    deg = Degrader(n)

    def one(i):
        return make_particle_fast(np.random.default_rng([seed, n, i]), sc, blobs, deg, max_tilt=max_tilt)

    with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        parts = list(ex.map(one, range(count)))
    return sc, parts
