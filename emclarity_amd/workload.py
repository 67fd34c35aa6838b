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


def local_grid():
    """The 1,045-orientation local search ([15,3,15,3], eulerSearch)."""
    return eulerSearch("C1", 15, 3, 15, 3)
