"""Seeded synthetic inputs shared by the CPU baseline, the tests and the
GPU paths (host numpy; data generation, not part of the hot path).

Reference volumes are sums of 3D Gaussian blobs evaluated analytically, so a
rotated/shifted copy is exact (the blob centres move; no interpolation).
A subtomogram is the reference rotated and shifted, multiplied in Fourier
space by a missing wedge about Y and a CTF, plus white noise at a given
SNR = var(signal) / var(noise).

Orientation convention: a particle generated with the Euler triple
(phi, theta, psi) is aligned by RotMat = BH_defineMatrix([phi, theta,
psi - phi], 'Bah', 'inv') (alignment/BH_alignRaw3d_v2.m:1175), i.e. the
particle is particle(x) = ref(RotMat^T (x - shift)).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .geometry import BH_defineMatrix


@dataclass
class Blobs:
    centres: np.ndarray   # (k, 3) in voxels relative to the box centre n//2 (x,y,z)
    sigma: np.ndarray     # (k,)
    amp: np.ndarray       # (k,)


def make_blobs(rng: np.random.Generator, n_blobs=40, radius=20.0, sigma=(2.0, 5.0), amp=(0.5, 1.0)):
    c = []
    while len(c) < n_blobs:
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) <= radius:
            c.append(p)
    return Blobs(np.array(c), rng.uniform(*sigma, n_blobs), rng.uniform(*amp, n_blobs))


def render(blobs: Blobs, n: int, R=None, shift=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Volume (n,n,n) (z,y,x) of the blobs moved to R c + shift about n//2."""
    R = np.eye(3) if R is None else np.asarray(R, dtype=np.float64)
    cen = blobs.centres @ R.T + np.asarray(shift, dtype=np.float64)
    ax = np.arange(n, dtype=np.float64) - n // 2
    vol = np.zeros((n, n, n))
    for (cx, cy, cz), s, a in zip(cen, blobs.sigma, blobs.amp):
        gx = np.exp(-((ax - cx) ** 2) / (2 * s * s))
        gy = np.exp(-((ax - cy) ** 2) / (2 * s * s))
        gz = np.exp(-((ax - cz) ** 2) / (2 * s * s))
        vol += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return vol


def fft_freqs(n: int):
    f = np.fft.fftfreq(n)
    return np.meshgrid(f, f, f, indexing="ij")  # (kz, ky, kx)


def wedge_mask(n: int, max_tilt_deg: float) -> np.ndarray:
    """Binary missing-wedge mask about Y in FFT order: sampled where the
    (kx, kz) direction lies within max_tilt of the kx axis."""
    kz, ky, kx = fft_freqs(n)
    t = math.tan(math.radians(max_tilt_deg))
    return (np.abs(kz) <= t * np.abs(kx) + 1e-12).astype(np.float64)


def ctf3d(n: int, pixel_A: float, defocus_um: float, kv=300.0, cs_mm=2.7, amp=0.1) -> np.ndarray:
    """Isotropic CTF -sin(chi) on the FFT-order grid (phase contrast sign as
    the reference's ctf kernel with sin(chi - amplitude phase))."""
    lam = 12.2643247 / math.sqrt(kv * 1e3 * (1.0 + kv * 1e3 * 0.978466e-6))  # Angstrom
    kz, ky, kx = fft_freqs(n)
    k2 = (kx * kx + ky * ky + kz * kz) / (pixel_A * pixel_A)
    df = defocus_um * 1e4
    chi = math.pi * 0.5 * cs_mm * 1e7 * lam ** 3 * k2 * k2 - math.pi * lam * df * k2 \
        - math.atan(amp / math.sqrt(1 - amp * amp))
    return np.sin(chi)


def degrade(vol, rng, max_tilt=60.0, pixel_A=2.5, defocus_um=3.0, snr=0.1, ctf=True):
    """Wedge x CTF in Fourier space, then white noise at the SNR. The CTF is
    applied phase-flipped (|CTF|), as in the CTF-corrected tomograms the
    alignment runs on: a raw sin(chi) inverts the low-frequency contrast and
    makes every cross-correlation against the reference negative."""
    n = vol.shape[0]
    F = np.fft.fftn(vol)
    W = wedge_mask(n, max_tilt) if max_tilt is not None else np.ones(F.shape)
    if ctf:
        W = W * np.abs(ctf3d(n, pixel_A, defocus_um))
    sig = np.real(np.fft.ifftn(F * W))
    if snr is None or snr <= 0:
        return sig
    noise = rng.standard_normal(sig.shape) * math.sqrt(sig.var() / snr)
    return sig + noise


def tilt_table(max_tilt=60.0, step=3.0, defocus_um=3.0, pixel_A=2.5, dose_per_tilt=3.0,
               wavelength_m=1.968697e-12, cs_m=2.7e-3, amp=0.1):
    """A synthetic 22-column TLT (the *_ctf.tlt layout, BH_weightMaskMex.m:12-30):
    col1 index, col4 tilt (deg), col11 exposure, col12 half-astig (m), col13
    astig angle (rad), col14 dose fraction, col15 -defocus (m), col16 pixel (m),
    col17 Cs (m), col18 wavelength (m), col19 amplitude contrast."""
    tilts = np.arange(-max_tilt, max_tilt + 1e-9, step)
    order = np.argsort(np.abs(tilts), kind="stable")
    T = np.zeros((len(tilts), 22))
    for i, t in enumerate(tilts):
        T[i, 0] = i + 1
        T[i, 3] = t
        T[i, 10] = dose_per_tilt * (np.where(order == i)[0][0] + 0.5)
        T[i, 13] = 1.0
        T[i, 14] = -defocus_um * 1e-6
        T[i, 15] = pixel_A * 1e-10
        T[i, 16] = cs_m
        T[i, 17] = wavelength_m
        T[i, 18] = amp
    return T


def sphere_mask(n: int, radius: float) -> np.ndarray:
    """EMC_maskShape('sphere', ..., 'kernel', false): 1 inside radius about n//2."""
    ax = np.arange(n) - n // 2
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    return ((x * x + y * y + z * z) <= radius * radius).astype(np.float32)


@dataclass
class AlignCase:
    ref: np.ndarray
    particle: np.ndarray
    euler: tuple          # ground-truth (phi, theta, psi)
    shift: np.ndarray     # ground-truth shift (x, y, z)
    rotmat: np.ndarray


def make_align_case(seed: int, n=64, euler=(30.0, 40.0, 50.0), shift=None, snr=0.1, n_blobs=40,
                    blob_radius=None, max_tilt=60.0, pixel_A=2.5, defocus_um=3.0, ctf=True):
    """Config-0-style particle: reference of n_blobs Gaussians, particle =
    reference rotated by the Euler triple and shifted U[-3,3] px, degraded."""
    rng = np.random.default_rng(seed)
    blobs = make_blobs(rng, n_blobs, radius=blob_radius or 0.3 * n)
    if shift is None:
        shift = rng.uniform(-3.0, 3.0, 3)
    R = BH_defineMatrix([euler[0], euler[1], euler[2] - euler[0]], "Bah", "inv")
    ref = render(blobs, n)
    clean = render(blobs, n, R, shift)
    part = degrade(clean, rng, max_tilt, pixel_A, defocus_um, snr, ctf)
    return AlignCase(ref.astype(np.float32), part.astype(np.float32), tuple(euler),
                     np.asarray(shift, dtype=np.float64), R)
