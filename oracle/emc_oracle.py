"""FP64 CPU restatement of emClarity's rotate-weight-correlate path.

TEST INFRASTRUCTURE ONLY. This module is the oracle: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker. The product (``emclarity_amd``) never
imports it and has no CPU fallback.

Every function restates one reference function in float64 numpy and cites
the reference file:line it follows. Volumes are numpy arrays of shape
``(nz, ny, nx)`` in C order, so the flat index is ``x + nx*(y + ny*z)``, the
same linear index as the reference's column-major MATLAB arrays
(``mexFiles/mexXform3d.cu:86``). Sizes are always given as ``(nx, ny, nz)``.

How each function is pinned (the tests in ``tests/test_oracle_*.py``):
  define_matrix        closed forms (orthonormal, det=+1, hand-computed 90 deg
                       cases, inv == forward^T)
  euler_search         orientation counts computed from the published grid
                       formula (15,244 / 4,600 / 43,784 / 1,045 / 23,281 / 1,236)
  xform3d              identity, integer shifts (exact translation), 90 deg
                       rotations (exact permutations about n//2), brute-force
                       trilinear on random 5^3/6^3 volumes, fwd/inv inverse pair
  ctf2d / sf3d         zeros of sin(chi) at analytic k, single 0-deg tilt ==
                       cosine slab x CTF^2, TLT fixtures from the reference docs
  bandpass3d           closed form at sampled radii
  band_limit_center_normalize   F(0)=0 and sum|F|^2 == numel^2
  xcf_translational    circular shift -> exact integer peak; symmetric peak COM
  xcf_rotational       CC(x,x)=1, Cauchy-Schwarz, scale invariance
  align / template search loops   tiny brute-force cases against ground truth
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as _sfft

# FP64 FFTs on every host core (pocketfft, as numpy.fft, multithreaded): the
# chunk-sized transforms of the template search (480^3) and the 2n-padded
# reference spectra stay within seconds.
def _rfftn(x):
    return _sfft.rfftn(x, workers=-1)


def _irfftn(x, s):
    return _sfft.irfftn(x, s=s, axes=(0, 1, 2), workers=-1)


def _ifftn(x):
    return _sfft.ifftn(x, workers=-1)

# --------------------------------------------------------------------------
# A3  BH_defineMatrix  (coordinates/BH_defineMatrix.m:51-126)
# --------------------------------------------------------------------------


def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _rx(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


def _ry(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def define_matrix(angles, convention="Bah", direction="forward"):
    """Rotation matrix from Euler angles in degrees.

    BH_defineMatrix.m:53 (deg->rad), :85-86 ('forward' negates), :87-102
    ('inv' reverses the order), :118-126 ('Bah' = Rz(a3) Rx(a2) Rz(a1)),
    :130-132 ('TILT', one angle about Y), :140-148 SPIDER, :155-163 Helical,
    :172-180 IMOD.
    """
    a = np.atleast_1d(np.asarray(angles, dtype=np.float64)) * (math.pi / 180.0)
    d = direction.lower()
    if d in ("forward", "invvector"):
        a = -a
    elif d in ("inv", "forwardvector"):
        a = a[::-1].copy()
    else:
        raise ValueError(f"Direction must be forward or inv, not {direction}")
    if convention == "Bah":
        return _rz(a[2]) @ _rx(a[1]) @ _rz(a[0])
    if convention == "TILT":
        c, s = math.cos(a[0]), math.sin(a[0])
        return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])
    if convention == "SPIDER":
        return _rz(a[2]) @ _ry(a[1]) @ _rz(a[0])
    if convention == "Helical":
        return _rz(a[2]) @ _rx(a[1]) @ _ry(a[0])
    if convention == "IMOD":
        return _rz(a[2]) @ _ry(a[1]) @ _rz(a[0])
    raise ValueError(f"Convention must be Bah,SPI,Helical, not {convention}")


# --------------------------------------------------------------------------
# A4  eulerSearch  (coordinates/eulerSearch.m:38-177)
# --------------------------------------------------------------------------


def matlab_colon(a, step, b):
    """MATLAB ``a:step:b`` element count and values (float)."""
    if step == 0 or (step > 0 and a > b) or (step < 0 and a < b):
        return np.zeros(0)
    n = int(math.floor((b - a) / step + 1e-10)) + 1
    return a + step * np.arange(n, dtype=np.float64)


@dataclass
class EulerGrid:
    theta: np.ndarray
    phi: list
    psi: np.ndarray
    theta_step: float
    psi_step: float
    n_at_theta: np.ndarray = field(default=None)

    def orientations(self):
        """Flat (phi, theta, psi) list in the reference loop order theta->phi->psi
        (BH_alignRaw3d_v2.m:1119-1170, BH_templateSearch3d_2.m:650-734)."""
        out = []
        for it, th in enumerate(self.theta):
            for ph in self.phi[it]:
                for ps in self.psi:
                    out.append((ph, th, ps))
        return np.array(out, dtype=np.float64).reshape(-1, 3)


def euler_search(symmetry, theta_max, theta_step, psi_max, psi_step, random_start=False):
    """Angular grid of eulerSearch (eulerSearch.m:38-124, limits :126-177).

    ``random_start`` must be False for parity: the reference draws it from an
    unseeded generator (eulerSearch.m:107-111).
    """
    bipolar = theta_max < 0
    theta_max = abs(theta_max)
    psi_max_full = 2.0 * psi_max                    # :58
    sym = symmetry[0]
    if sym == "C":
        nsym = int(symmetry[1:])
        psi_max_full = min(psi_max_full, 360.0 / nsym)
        theta_max = min(180.0, theta_max)
        phi_max = 360.0
    elif sym == "D":
        nsym = int(symmetry[1:])
        psi_max_full = 360.0 / nsym                  # :143 min() of one value
        theta_max = min(theta_max, 90.0)
        phi_max = 360.0
    elif sym == "O":
        psi_max_full = min(psi_max_full, 90.0)
        theta_max = min(theta_max, 54.7)
        phi_max = 90.0
    elif sym == "I":
        psi_max_full = min(psi_max_full, 180.0)
        theta_max = 31.7
        phi_max = 180.0
    else:
        raise ValueError(symmetry)
    psi = matlab_colon(-psi_max_full / 2.0, psi_step, psi_max_full / 2.0)  # :73
    theta = matlab_colon(0.0, theta_step, theta_max)                        # :76
    if theta.size == 0:
        theta = np.zeros(1)
    if bipolar:
        theta = np.concatenate([theta, (180.0 - theta)[::-1]])
    phis = []
    for th in theta:                                                        # :92-116
        if th == 0.0 or th == 180.0:
            phi_step = phi_max
        else:
            phi_step = abs(theta_step / math.sin(math.radians(th)))
            if phi_step > phi_max:
                phi_step = phi_max
            else:
                phi_step = phi_max / math.floor(phi_max / phi_step + 0.5)
        phis.append(matlab_colon(0.0, phi_step, phi_max - 1.0))
    if psi.size == 0:
        psi = np.zeros(1)
    n_at = np.array([len(p) * len(psi) for p in phis])
    return EulerGrid(theta, phis, psi, theta_step, psi_step, n_at)


# --------------------------------------------------------------------------
# A1/A2  mexXform3d + interpolator  (mexFiles/mexXform3d.cu:45-86,160-172,
#        293-296; testScripts/interpolator.m:99-171)
# --------------------------------------------------------------------------


def _trilinear_zero(vol, px, py, pz):
    """Trilinear sample of vol at index coordinates, zero outside [0,n-1]
    (border addressing, mexXform3d.cu:269-271)."""
    nz, ny, nx = vol.shape
    x0 = np.floor(px).astype(np.int64)
    y0 = np.floor(py).astype(np.int64)
    z0 = np.floor(pz).astype(np.int64)
    fx, fy, fz = px - x0, py - y0, pz - z0
    out = np.zeros(px.shape, dtype=np.float64)
    for dz in (0, 1):
        wz = fz if dz else 1.0 - fz
        zi = z0 + dz
        okz = (zi >= 0) & (zi < nz)
        for dy in (0, 1):
            wy = fy if dy else 1.0 - fy
            yi = y0 + dy
            oky = okz & (yi >= 0) & (yi < ny)
            for dx in (0, 1):
                wx = fx if dx else 1.0 - fx
                xi = x0 + dx
                ok = oky & (xi >= 0) & (xi < nx)
                v = np.zeros(px.shape, dtype=np.float64)
                v[ok] = vol[zi[ok], yi[ok], xi[ok]]
                out += wx * wy * wz * v
    return out


def xform3d(vol, M, t, fwd):
    """Resample ``vol`` by the 3x3 matrix ``M`` (the matrix the caller passes,
    read row-wise by the kernel: mexXform3d.cu:170-172) and shift ``t``.

    fwd:  p = M (u - t)   (shift negated :160-164, added before rotation :49-54)
    inv:  p = M u + t     (:59-64)
    with u = index - n//2. The texture sample point in index units is
    p + n//2 for both parities (normalised coords + the even-size half texel,
    :68-74, :293-296); trilinear, zero outside.
    """
    vol = np.asarray(vol, dtype=np.float64)
    nz, ny, nx = vol.shape
    M = np.asarray(M, dtype=np.float64).reshape(3, 3)
    t = np.asarray(t, dtype=np.float64).reshape(3)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u = np.stack([x - nx // 2, y - ny // 2, z - nz // 2]).astype(np.float64)
    if fwd:
        u = u - t[:, None, None, None]
        p = np.einsum("ij,j...->i...", M, u)
    else:
        p = np.einsum("ij,j...->i...", M, u) + t[:, None, None, None]
    return _trilinear_zero(vol, p[0] + nx // 2, p[1] + ny // 2, p[2] + nz // 2)


def _frame(a, b):
    e1 = a / np.linalg.norm(a)
    e2 = b - (b @ e1) * e1
    e2 /= np.linalg.norm(e2)
    return np.stack([e1, e2, np.cross(e1, e2)], axis=1)


def _vertex_group(vertices):
    """Every rotation of a regular polyhedron, by enumeration: the rotation
    taking a fixed vertex and one of its nearest neighbours onto each vertex
    and each of its nearest neighbours."""
    V = np.asarray(vertices, dtype=np.float64)
    d = np.linalg.norm(V[:, None] - V[None], axis=2)
    near = d[0][d[0] > 1e-9].min()
    nb = lambda i: [j for j in range(len(V)) if abs(d[i, j] - near) < 1e-9]
    F0 = _frame(V[0], V[nb(0)[0]])
    return [_frame(V[a], V[b]) @ F0.T for a in range(len(V)) for b in nb(a)]


def symmetry_matrices(symmetry, convention="Bah"):
    """Symmetry operators of interpolator.m:173-380. C<n>: BH_defineMatrix([0,
    0, i 360/n]) (:188-198). D<n>: the reference makes 2n matrices but keeps
    nSymMats = n (:199-210), so its interpolator uses the n C_n ones ('Bah').
    O: the octahedron (+-x, +-y, +-z)'s 24 rotations. I / I2: the 60
    rotations of the icosahedron with vertices at the cyclic permutations of
    (0, +-1, +-phi) / (+-1, 0, +-phi), the two settings of :245-372."""
    s = symmetry.strip()
    if s[0] in "CD":
        n = int(s[1:])
        conv = convention if s[0] == "C" else "Bah"
        return [define_matrix([0.0, 0.0, i * 360.0 / n], conv, "forward") for i in range(n)]
    if s == "O":
        return _vertex_group([v for a in range(3) for v in (np.eye(3)[a], -np.eye(3)[a])])
    if s in ("I", "I1", "I2"):
        phi = (1.0 + math.sqrt(5.0)) / 2.0
        base = (1.0, 0.0, phi) if s == "I2" else (0.0, 1.0, phi)
        nz = [i for i in range(3) if base[i] != 0]
        V = []
        for sa in (1, -1):
            for sb in (1, -1):
                v = np.array(base)
                v[nz[0]] *= sa
                v[nz[1]] *= sb
                V += [np.roll(v, k) for k in range(3)]
        return _vertex_group(V)
    raise NotImplementedError(symmetry)


def _as_matrix(a):
    """A 3x3 matrix as given, or a 9-vector read column-major, as MATLAB's
    reshape(angles, 3, 3) (interpolator.m:163) reads geometry columns 17:25."""
    a = np.asarray(a, dtype=np.float64)
    return a.reshape(3, 3) if a.ndim == 2 else a.reshape(3, 3, order="F")


def interp3d(vol, angles, shifts, convention="Bah", direction="forward", symmetry="C1"):
    """interpolator.interp3d (interpolator.m:99-154): Euler triple -> matrix
    (:159-161), sum over symmetry mates M*S_i (:108-143), / nSym (:145-147)."""
    a = np.asarray(angles, dtype=np.float64)
    M = define_matrix(a, convention, direction) if a.size == 3 else _as_matrix(a)
    fwd = direction.lower() == "forward"
    mats = symmetry_matrices(symmetry, convention)
    out = None
    for S in mats:
        r = xform3d(vol, M @ S, shifts, fwd)
        out = r if out is None else out + r
    return out / len(mats) if len(mats) > 1 else out


# --------------------------------------------------------------------------
# A5/A7  CTF (mexFiles/utils/ctf.cu:12-128; include/core_headers.cuh:40-75)
# --------------------------------------------------------------------------

_EXP_A, _EXP_B, _EXP_C = 0.24499, -0.8325, 2.81410    # ctf.cu:5-8


@dataclass
class CtfParams:
    """ctfParams constructor (core_headers.cuh:63-75)."""
    pixel: float
    wavelength: float
    cs_mm: float
    amp: float
    df1: float
    df2: float
    astig_deg: float

    @property
    def amp_phase(self):
        return math.atan(self.amp / math.sqrt(1.0 - self.amp ** 2))

    @property
    def dmean(self):
        return 0.5 * (self.df1 + self.df2)

    @property
    def ddiff(self):
        return 0.5 * (self.df1 - self.df2)

    @property
    def astig(self):
        a = self.astig_deg
        return a * math.pi / 180.0 - math.pi * float(np.rint(a / 180.0))

    @property
    def cs_term(self):
        return math.pi * 0.5 * self.cs_mm * 1e7 * self.wavelength ** 3

    @property
    def df_term(self):
        return math.pi * self.wavelength


def ctf2d(nx, ny, p: CtfParams, half=False, sq=False, centered=True,
          weighted=False, occupancy=1.0, exposure=0.0, spacing_from_full_nx=True):
    """2D CTF image, shape (ny, nxh) with nxh = nx//2+1 for half grids
    (kernels ctf.cu:12-65 plain, :68-128 weighted).

    The Fourier voxel size along x is 1/(pixel*nx) in the mexCTF gateway
    (mexCTF.cu:44, full nx) and 1/(pixel*ctf_dims.x) in mexSF3D
    (mexSF3D.cu:148, :234-235, after the half-grid reduction);
    ``spacing_from_full_nx`` selects between them (equal for full grids)."""
    cx = nx // 2 + 1 if half else nx
    o_x = 0 if half else cx // 2
    o_y = ny // 2
    fvx = 1.0 / (p.pixel * (nx if spacing_from_full_nx else cx))
    fvy = 1.0 / (p.pixel * ny)
    y, x = np.meshgrid(np.arange(ny), np.arange(cx), indexing="ij")
    if centered:
        y = y - o_y
        if not half:
            x = x - o_x
    else:
        y = np.where(y > o_y, y - ny, y)
        if not half:
            x = np.where(x > o_x, x - cx, x)
    ky, kx = y * fvy, x * fvx
    phi = np.arctan2(ky, kx)
    r2 = kx * kx + ky * ky
    chi = p.cs_term * r2 ** 2 - p.df_term * r2 * (p.dmean + p.ddiff * np.cos(2.0 * (phi - p.astig))) - p.amp_phase
    v = np.sin(chi)
    if sq:
        v = v * v
    if weighted:
        rw = occupancy * np.where(x == 0, 0.2, np.abs(x).astype(np.float64))
        with np.errstate(divide="ignore"):
            pw = np.where(r2 > 0, np.power(np.where(r2 > 0, r2, 1.0), _EXP_B), np.inf)
        v = v * rw * np.exp((-0.5 * exposure) / (1.0 * (_EXP_A * pw + _EXP_C)))
    return v


def _tex2d_linear_border(img, tu, tv):
    """tex2D with normalised coords, linear filter, border 0 (texel centre
    convention x_tex = tu*W - 0.5)."""
    H, W = img.shape
    xt = tu * W - 0.5
    yt = tv * H - 0.5
    x0 = np.floor(xt).astype(np.int64)
    y0 = np.floor(yt).astype(np.int64)
    fx, fy = xt - x0, yt - y0
    out = np.zeros(np.broadcast(xt, yt).shape)
    for dy in (0, 1):
        wy = fy if dy else 1.0 - fy
        yi = y0 + dy
        for dx in (0, 1):
            wx = fx if dx else 1.0 - fx
            xi = x0 + dx
            ok = (xi >= 0) & (xi < W) & (yi >= 0) & (yi < H)
            v = np.zeros(out.shape)
            xb, yb = np.broadcast_arrays(xi, yi)
            okb = np.broadcast_to(ok, out.shape)
            v[okb] = img[yb[okb], xb[okb]]
            out += wx * wy * v
    return out


def sf3d(size, tilts, half=False, sq=True):
    """mexSF3D: per-tilt weighted 2D CTF (ctf.cu:68-128) back-projected as a
    cosine slab of half-thickness 4 px about the tilted plane and summed over
    tilts (mexSF3D.cu:25-79 kernel, :125-277 gateway).

    ``tilts`` is a list of dicts with keys pixel, wavelength, cs, df1, df2,
    astig, amp, tilt, exposure, occupancy (one per tilt, the gateway's
    per-tilt arrays :104-114). Returns (nz, ny, nx).
    """
    nx, ny, nz = (int(s) for s in size)
    out = np.zeros((nz, ny, nx))
    ssx = 0.5 if nx % 2 == 0 else 0.0
    ssy = 0.5 if ny % 2 == 0 else 0.0
    ssz = 0.5 if nz % 2 == 0 else 0.0
    cx = nx // 2 + 1 if half else nx
    for t in tilts:
        p = CtfParams(t["pixel"], t["wavelength"], t["cs"], t["amp"], t["df1"], t["df2"], t["astig"])
        img = ctf2d(nx, ny, p, half=half, sq=sq, centered=True, weighted=True,
                    occupancy=t["occupancy"], exposure=t["exposure"], spacing_from_full_nx=False)
        s = math.sin(math.radians(-t["tilt"]))                              # :134
        c = math.cos(math.radians(-t["tilt"]))
        x = np.arange(nx, dtype=np.float64)
        y = np.arange(ny, dtype=np.float64)
        z = np.arange(nz, dtype=np.float64)
        u = x - nx / 2.0                                                    # :47 float division
        tw_pre = -u * s + ssz                                               # :48
        tv = (y - (ny // 2) + ssy) / ny + 0.5                               # :49
        w = z - (nz // 2)                                                   # :53
        tw = tw_pre[None, :] + w[:, None] * c                               # (nz, nx) :54
        keep = ~((tw < -4.0) | (tw > 4.0))                                  # :56
        zw = (0.5 + 0.5 * np.cos(tw * (math.pi / 4.0))) * 0.25              # :62
        tu = (u[None, :] * c + w[:, None] * s + ssx) / nx + 0.5             # :64-72
        # sample the CTF texture for every (z, y, x)
        samp = _tex2d_linear_border(img, tu[:, None, :], tv[None, :, None])
        out += np.where(keep[:, None, :], zw[:, None, :] * samp, 0.0)
    return out


def tlt_to_tilts(TLT, sampling, xyz_subtomo, recon_geometry, size):
    """BH_weightMaskMex marshalling (masking/BH_weightMaskMex.m:7-51): TLT rows
    (22 columns) -> per-tilt SF3D arguments, with the particle-height defocus
    offset (:33-37) and occupancy = dose fraction x elastic fraction (:41-44)."""
    TLT = np.asarray(TLT, dtype=np.float64)
    TLT = TLT[np.argsort(TLT[:, 0], kind="stable")]                         # :12 sortrows(TLT,1)
    tilt = TLT[:, 3]
    pixel = TLT[:, 15] * sampling * 1e10
    recon_shift = np.asarray(recon_geometry[1], dtype=np.float64)
    origin = np.ceil((np.asarray(recon_geometry[0][:3], dtype=np.float64) + 1.0) / 2.0)
    prj = np.asarray(xyz_subtomo, dtype=np.float64) - origin + recon_shift
    cs = TLT[:, 16] * 1e3
    wl = TLT[:, 17] * 1e10
    phase = TLT[:, 18]
    defocus = -TLT[:, 14] * 1e10
    ddf = TLT[:, 11] * 1e10
    dphi = TLT[:, 12] * 180.0 / math.pi
    exposure = TLT[:, 10]
    for i in range(TLT.shape[0]):
        r = define_matrix([TLT[i, 3]], "TILT", "forwardVector")
        pc = r @ prj
        defocus[i] = defocus[i] - pc[2] * pixel[i]
    dose = TLT[:, 13] / TLT[:, 13].mean()
    elastic = np.exp(-75.0 / (np.cos(np.radians(TLT[:, 3])) * 400.0))
    elastic = elastic / elastic.max()
    occ = elastic * dose
    return [dict(pixel=pixel[i], wavelength=wl[i], cs=cs[i], df1=defocus[i] + ddf[i],
                 df2=defocus[i] - ddf[i], astig=dphi[i], amp=phase[i], tilt=tilt[i],
                 exposure=exposure[i], occupancy=occ[i]) for i in range(TLT.shape[0])]


def weight_mask_mex(size, sampling, TLT, xyz_subtomo, recon_geometry):
    """BH_weightMaskMex (:1-59): SF3D then shift so min == 1e-4 (:55)."""
    w = sf3d(size, tlt_to_tilts(TLT, sampling, xyz_subtomo, recon_geometry, size), half=False, sq=True)
    return w - (w.min() - 1e-4)


# --------------------------------------------------------------------------
# A10 BH_bandpass3d (masking/BH_bandpass3d.m:54-170) and A20 radial grid
# --------------------------------------------------------------------------


def fft_radius(size):
    """Radial frequency grid in FFT order, cycles/pixel
    (BH_multi_gridCoordinates.m:209-215, :239-243, :303)."""
    nx, ny, nz = size
    def ax(n):
        return np.concatenate([np.arange(0, n // 2 + 1), -np.arange((n - 1) // 2, 0, -1)]) / n
    z, y, x = np.meshgrid(ax(nz), ax(ny), ax(nx), indexing="ij")
    return np.sqrt(x * x + y * y + z * z)


def bandpass3d(size, high_thresh, high_cut, low_cut, pixel):
    """BH_bandpass3d for a numeric pixel size (:62-64, :102-115, :124-170).
    Returns (nz, ny, nx) in FFT order."""
    size = tuple(int(s) for s in size)
    if not (0 <= high_thresh < 1):
        raise ValueError("HIGH_THRESH must be between 0 and 1")
    hc = pixel / high_cut if high_cut != 0 else 0.0
    lc = pixel / low_cut
    if lc > 0:
        lroll = 2.0 / (min(size[:2]) if size[2] == 1 else min(size))
    else:
        lc = abs(lc)
        lroll = math.sqrt(-((0.5 - lc) ** 2) / (2.0 * math.log(1e-3)))
    hroll = math.sqrt(-(hc ** 2) / (2.0 * math.log(high_thresh))) if high_cut != 0 else 0.0
    r = fft_radius(size)
    g = lambda x, m, s: np.exp(-((x - m) ** 2) / (2.0 * s * s))
    w = np.where(r < lc, 1.0, g(r, lc, lroll))
    if hc != 0:
        w = np.where(r <= hc, g(r, hc, hroll), w)
    return w


# --------------------------------------------------------------------------
# A12 BH_padZeros3d (masking/BH_padZeros3d.m:37-229), A17 padVal
# --------------------------------------------------------------------------


def pad_val(size1, size2):
    """BH_multi_padVal (:7-22): rows [fwdLow; fwdHigh; invLow; invHigh]."""
    s1 = np.asarray(size1, dtype=np.int64)
    s2 = np.asarray(size2, dtype=np.int64)
    d = s2 - s1
    od, oi = d % 2, s1 % 2
    fwd = np.stack([d // 2 + oi * od, d // 2 + (1 - oi) * od])
    d = s1 - s2
    od, oi = d % 2, s2 % 2
    inv = np.stack([d // 2 + oi * od, d // 2 + (1 - oi) * od])
    return np.concatenate([fwd, inv])


_TAPER = 0.5 + 0.5 * np.cos(np.arange(7) * math.pi / 7.0)        # :110-117


def pad_zeros3d(img, pad_low, pad_top, extrap_val=0.0, taper=False):
    """Pad/trim an (nz,ny,nx) volume; pads are given as (x,y,z).
    Negative = trim (:79-93); constant fill value (:59-69, :154-156);
    'singleTaper' 7-px cosine edge (:189-205)."""
    img = np.array(img, dtype=np.float64, copy=True)
    pl = np.asarray(pad_low, dtype=np.int64)
    pt = np.asarray(pad_top, dtype=np.int64)
    tl = np.where(pl < 0, -pl, 0)
    tt = np.where(pt < 0, -pt, 0)
    pl = np.where(pl >= 0, pl, 0)
    pt = np.where(pt >= 0, pt, 0)
    nz, ny, nx = img.shape
    img = img[tl[2]:nz - tt[2], tl[1]:ny - tt[1], tl[0]:nx - tt[0]].copy()
    if taper:
        ft = _TAPER[::-1]
        ev = extrap_val
        # order follows :190-205: y-low, x-low, z-low, y-high, x-high, z-high
        img[:, :7, :] = img[:, :7, :] * ft[None, :, None] + (ev * (1 - _TAPER))[::-1][None, :, None]
        img[:, :, :7] = img[:, :, :7] * ft[None, None, :] + (ev * (1 - _TAPER))[::-1][None, None, :]
        img[:7, :, :] = img[:7, :, :] * ft[:, None, None] + (ev * (1 - _TAPER))[::-1][:, None, None]
        img[:, -7:, :] = img[:, -7:, :] * _TAPER[None, :, None] + (ev * (1 - _TAPER))[None, :, None]
        img[:, :, -7:] = img[:, :, -7:] * _TAPER[None, None, :] + (ev * (1 - _TAPER))[None, None, :]
        img[-7:, :, :] = img[-7:, :, :] * _TAPER[:, None, None] + (ev * (1 - _TAPER))[:, None, None]
    nz, ny, nx = img.shape
    out = np.full((nz + pl[2] + pt[2], ny + pl[1] + pt[1], nx + pl[0] + pt[0]), float(extrap_val))
    out[pl[2]:pl[2] + nz, pl[1]:pl[1] + ny, pl[0]:pl[0] + nx] = img
    return out


# --------------------------------------------------------------------------
# A11 BH_bandLimitCenterNormalize (masking/BH_bandLimitCenterNormalize.m:27-61)
# --------------------------------------------------------------------------


def band_limit_center_normalize(img, bandpass, pad):
    """pad with mean(img) (:48) -> fftn -> F(0)=0 (:51) -> x bandpass (:55)
    -> F / (sqrt(sum|F|^2)/numel) (:61). ``bandpass`` None means none;
    ``pad`` is [[lowx,lowy,lowz],[topx,topy,topz]]."""
    img = np.asarray(img, dtype=np.float64)
    p = pad_zeros3d(img, pad[0], pad[1], extrap_val=float(img.mean()))
    F = np.fft.fftn(p)
    F.flat[0] = 0.0
    if bandpass is not None:
        F = F * bandpass
    return F / (math.sqrt(float(np.sum(np.abs(F) ** 2))) / F.size)


# --------------------------------------------------------------------------
# A13 BH_multi_xcf_Translational (statistics/BH_multi_xcf_Translational.m:31-103)
# --------------------------------------------------------------------------


def matlab_round(x):
    return np.where(np.asarray(x) >= 0, np.floor(np.asarray(x) + 0.5), np.ceil(np.asarray(x) - 0.5))


def xcf_translational(A, B, peak_mask, peak_com, return_map=False):
    """Returns the 1x3 shift (x,y,z) in pixels. ``peak_mask`` None = no mask."""
    M = np.fft.fftshift(np.real(np.fft.ifftn(A * B)))                  # :31
    min_val = M.min()                                                   # :33
    if peak_mask is not None:
        M = np.where(np.asarray(peak_mask) < 0.95, min_val, M)          # :34
    nz, ny, nx = M.shape
    idx = int(np.argmax(M.ravel()))                                     # :40 first max
    max_val = M.ravel()[idx]
    z, rem = divmod(idx, nx * ny)
    y, x = divmod(rem, nx)
    peak = np.array([x + 1, y + 1, z + 1], dtype=np.float64)           # :41-42 1-based
    com = np.asarray(peak_com, dtype=np.int64).reshape(3)
    if (not np.isnan(max_val)) and max_val != 0:                        # :45
        for icom in (1, 2):                                             # :47
            if np.all(com != 0):                                        # :49
                pc = peak if icom == 1 else matlab_round(peak)          # :50-55
                lo = (pc - com).astype(np.int64)
                hi = (pc + com).astype(np.int64)
                if np.any(lo < 1) or np.any(hi > np.array([nx, ny, nz])):
                    raise IndexError("failed to box out COM calc")       # :64-81
                box = M[lo[2] - 1:hi[2], lo[1] - 1:hi[1], lo[0] - 1:hi[0]] - min_val
                gz, gy, gx = np.meshgrid(np.arange(-com[2], com[2] + 1), np.arange(-com[1], com[1] + 1),
                                         np.arange(-com[0], com[0] + 1), indexing="ij")
                s = box.sum()
                cm = np.array([(box * gx).sum(), (box * gy).sum(), (box * gz).sum()]) / s   # :84-86
                peak = pc + cm                                          # :88
    else:
        peak = np.zeros(3)                                              # :95
    peak = peak - np.ceil((np.array([nx, ny, nz]) + 1) / 2.0)           # :97
    peak = np.where(np.isnan(peak), 0.0, peak)                          # :99-103
    if return_map:
        return peak, M, max_val
    return peak


def fwd_fft(vol, doNorm=0, doCenter=False, bp_half=None, ctf=None):
    """fourierTransformer.fwdFFT (testScripts/fourierTransformer.m:92-145) on
    a (nz, ny, nx) volume: R2C half spectrum along x (the last axis here,
    MATLAB's first), x ctf, x (1/sqrt(numel))^doNorm when not centring
    (:130-132), x bandpass (:133-136), and with centring F(1) = 0 then
    F / sqrt(2 sum|F(1:end-invTrim,:,:)|^2) (:137-140), invTrim = 2 for
    even nx, 1 for odd (:62-66)."""
    vol = np.asarray(vol, np.float64)
    nx = vol.shape[-1]
    ft = np.fft.rfftn(vol)
    if ctf is not None:
        ft = ft * ctf
    if doNorm and not doCenter:
        ft = ft * (1.0 / math.sqrt(vol.size)) ** doNorm
    if bp_half is not None:
        ft = ft * bp_half
    if doCenter:
        ft = ft.copy()
        ft.flat[0] = 0
        inv_trim = 1 if nx % 2 else 2
        ft = ft / math.sqrt(2.0 * np.sum(np.abs(ft[..., :ft.shape[-1] - inv_trim]) ** 2))
    return ft


# --------------------------------------------------------------------------
# A14 BH_multi_xcf_Rotational (statistics/BH_multi_xcf_Rotational.m:29-80)
# --------------------------------------------------------------------------


def xcf_rotational(P, R, Wp, Wr, wCCC=0):
    """CC = sum Re(Wr P R Wp) / sqrt(sum|R Wp|^2 * sum|P Wr|^2) (:73-78);
    NaN first element of a weight -> weight 1 (:34-40).

    wCCC a list [abs flag, w, shell_1, shell_2, ...] (the reference's cell,
    :50-70): each shell a list of 1-based MATLAB linear indices; volumes are
    (nz, ny, nx), whose flat C order runs x fastest like MATLAB's linear
    index, so index i names flat element i-1. Per shell the CC above over
    its voxels only, c; the result is sum over shells of
    (c^2 / (c + w^-1.5))^3, its absolute value when the flag is set (:65-69).
    """
    Wr = np.asarray(Wr, dtype=np.float64)
    Wp = np.asarray(Wp, dtype=np.float64)
    if np.isnan(Wr.flat[0]):
        Wr = np.float64(1.0)
    if np.isnan(Wp.flat[0]):
        Wp = np.float64(1.0)
    if isinstance(wCCC, (list, tuple)):
        shape = np.shape(P)
        P = np.asarray(P, np.complex128).ravel()
        R = np.asarray(R, np.complex128).ravel()
        Wp = np.broadcast_to(Wp, shape).ravel()
        Wr = np.broadcast_to(Wr, shape).ravel()
        take_abs = bool(np.asarray(wCCC[0]).ravel()[0])
        k = float(np.asarray(wCCC[1]).ravel()[0]) ** -1.5
        total = 0.0
        for shell in wCCC[2:]:
            j = np.asarray(shell, np.int64).ravel() - 1
            num = np.sum(np.real(Wr[j] * P[j] * R[j] * Wp[j]))
            den = np.sum(np.abs(R[j] * Wp[j]) ** 2) * np.sum(np.abs(P[j] * Wr[j]) ** 2)
            c = float(num / math.sqrt(den)) if den > 0 else float("nan")
            t = (c * c / (c + k)) ** 3
            total += abs(t) if take_abs else t
        return total
    num = np.sum(np.real(Wr * P * R * Wp))
    den = np.sum(np.abs(R * Wp) ** 2) * np.sum(np.abs(P * Wr) ** 2)
    return float(num / math.sqrt(den))


# --------------------------------------------------------------------------
# A15 the alignRaw per-particle loop (alignment/BH_alignRaw3d_v2.m:1119-1765)
# --------------------------------------------------------------------------


@dataclass
class AlignInputs:
    """Everything the per-particle loop reads (BH_alignRaw3d_v2.m:800-1110).
    Volumes are (nz,ny,nx). W = window edge, C = calc edge."""
    particle: np.ndarray            # iparticle after padZeros3d 'singleTaper' (W^3)
    ref: np.ndarray                 # ref_FT2 (untrimmed reference, W^3)   :1108
    ref_wgt_rot: np.ndarray         # refWgtROT (C^3 centred)              :1109
    ref_wgt: object                 # refWGT (C^3 FFT order) or scalar     :1385
    ref_ft1: np.ndarray             # conj(bandLimitCenterNormalize(ref*volMask,1,...)) :607
    vol_mask: np.ndarray            # volMask (W^3)
    peak_mask: np.ndarray           # peakMask (C^3)
    bandpass: np.ndarray            # bandpassFilt (C^3 FFT order)
    wedge: np.ndarray               # SF3D from BH_weightMaskMex (C^3 centred) :1093
    angles: np.ndarray              # particle's current 3x3 matrix (col-major 9 in geometry)
    shift_val: np.ndarray           # BH_isWindowValid SHIFTS (:1042)
    pad_calc: np.ndarray            # [[low],[top]] W->C
    peak_com: tuple = (3, 3, 3)
    symmetry: str = "C1"
    # stage-1 particle bandpass when it is another reference's (several
    # references: stage 1 runs with the first of refToAlign, :1278-1318)
    bandpass_stage1: object = None
    # wCCC: 0 (plain CCC) or the shell cell [abs flag, w, shells...] of
    # BH_multi_xcf_Rotational (:1355-1359)
    wccc: object = 0


def ref_ft1_of(ref, vol_mask, pad_calc):
    """ref_FT1 = conj(bandLimitCenterNormalize(ref.*volMask, 1, ...)) (:607-608)."""
    return np.conj(band_limit_center_normalize(ref * vol_mask, None, pad_calc))


def align_orientation(inp: AlignInputs, rot_mat, est_peak0):
    """One orientation of the two-stage body (:1182-1402). Returns
    (ccc, est_peak) where est_peak is the stage-1 shift + est_peak0."""
    R = rot_mat
    # stage 1: rotate the reference (:1287-1336)
    trim = inp.particle                                    # padWindow == 0
    i_rot_ref = xform3d(inp.ref, R.T, est_peak0, True)
    if inp.ref_wgt_rot is None:                            # no reference weight: 1
        i_rot_wdg = np.ones(inp.peak_mask.shape)
    else:
        i_rot_wdg = xform3d(inp.ref_wgt_rot, R.T, [0, 0, 0], True)
    i_rot_mask = xform3d(inp.peak_mask, R.T, [0, 0, 0], True)
    i_rot_ref = band_limit_center_normalize(i_rot_ref, None, inp.pad_calc)
    bp1 = inp.bandpass if inp.bandpass_stage1 is None else inp.bandpass_stage1
    rot_part = band_limit_center_normalize(trim, bp1, inp.pad_calc)
    wedge_ifft = np.fft.ifftshift(inp.wedge)
    peak = xcf_translational(rot_part * np.fft.ifftshift(i_rot_wdg), np.conj(i_rot_ref) * wedge_ifft,
                             i_rot_mask, inp.peak_com)
    est = peak + np.asarray(est_peak0, dtype=np.float64)   # :1341-1344
    # stage 2: rotate the particle (:1199-1256, :1375-1387)
    part = interp3d(inp.particle, R, est, "Bah", "inv", inp.symmetry)
    wmask = interp3d(inp.wedge, R, [0, 0, 0], "Bah", "inv", inp.symmetry)
    rot_part2 = band_limit_center_normalize(part * inp.vol_mask, inp.bandpass, inp.pad_calc)
    ccc = xcf_rotational(rot_part2, inp.ref_ft1, np.fft.ifftshift(wmask), inp.ref_wgt, inp.wccc)
    return ccc, est


def align_initial(inp: AlignInputs):
    """getInitialCCC (:1222-1232, :1355-1369): CCC at the current pose."""
    A = _as_matrix(inp.angles)
    part = interp3d(inp.particle, A, inp.shift_val, "Bah", "inv", inp.symmetry)
    wmask = interp3d(inp.wedge, A, [0, 0, 0], "Bah", "inv", inp.symmetry)
    rp = band_limit_center_normalize(part * inp.vol_mask, inp.bandpass, inp.pad_calc)
    return xcf_rotational(rp, inp.ref_ft1, np.fft.ifftshift(wmask), inp.ref_wgt, inp.wccc)


def align_grid(inp: AlignInputs, grid_angles):
    """Coarse grid (:1119-1419) over an explicit (phi, theta, psi) list with
    the azimuthal randomiser zeroed. Returns (ccc[n], shifts[n,3])."""
    A = _as_matrix(inp.angles)
    n = len(grid_angles)
    ccc = np.zeros(n)
    est = np.zeros((n, 3))
    for i, (phi, theta, psi) in enumerate(grid_angles):
        R = A @ define_matrix([phi, theta, psi - phi], "Bah", "inv")   # :1175-1176
        ccc[i], est[i] = align_orientation(inp, R, inp.shift_val)
    return ccc, est


def sortrows_desc_first(values):
    """Index of the row sortrows(...,-6) puts first: max value, earliest on
    ties (MATLAB sortrows is stable)."""
    v = np.asarray(values)
    return int(np.flatnonzero(v == v.max())[0])


# --------------------------------------------------------------------------
# A16 template search (alignment/BH_templateSearch3d_2.m:632-982)
# --------------------------------------------------------------------------


def swap_quadrants_half(size):
    """exp(-2 pi i (dU (nx//2+1) + dV (ny//2+1) + dW (nz//2+1))) on the half
    grid (:632-641), dU etc. centred frequency grids (origin at n//2)."""
    nx, ny, nz = size
    def ax(n):
        return np.arange(-(n // 2), (n - 1) // 2 + 1) / n
    z, y, x = np.meshgrid(ax(nz), ax(ny), ax(nx), indexing="ij")
    sq = np.exp(-2j * np.pi * (x * (nx // 2 + 1) + y * (ny // 2 + 1) + z * (nz // 2 + 1)))
    return sq[:, :, : nx // 2 + 1]


def matlab_std(x):
    x = np.asarray(x, dtype=np.float64).ravel()
    return float(np.std(x, ddof=1))


def template_search_chunk(tomo_chunk, template, grid_angles, pad_bin, trim_valid,
                          prev_peak=None, prev_angle=None, first_angle_id=1, init_from_first=True,
                          return_margin=False):
    """Per chunk, per orientation (:713-887): tempRot = interp3d(template,
    [phi,theta,psi-phi], [1,1,1], 'forward') (:744-747), zero-pad into the
    chunk (:751-754), ccf = trim(real(C2R(swapQ R2C(chunk) conj(R2C(tempPAD)))))
    (:759-761), / std (:764), strict-< max/argmax update (:851-854, :883-887).

    Returns (peak, angle_id) over the valid area, angle ids 1-based. With
    return_margin, also the per-voxel margin between the best and the
    second-best score of this call's orientations (and the previous peak,
    when given): where it exceeds twice a score's error, the argmax id is
    determined by the scores alone."""
    nz, ny, nx = tomo_chunk.shape
    size = (nx, ny, nz)
    tomo_fou = swap_quadrants_half(size) * _rfftn(np.asarray(tomo_chunk, dtype=np.float64))
    mag = None if prev_peak is None else prev_peak.copy()
    ang = None if prev_angle is None else prev_angle.copy()
    second = None if prev_peak is None else np.full(prev_peak.shape, -np.inf)
    aid = first_angle_id
    for (phi, theta, psi) in grid_angles:
        trot = interp3d(template, [phi, theta, psi - phi], [1, 1, 1], "Bah", "forward")
        tpad = pad_zeros3d(trot, pad_bin[0], pad_bin[1])
        temp_fou = np.conj(_rfftn(tpad))
        ccf = _irfftn(tomo_fou * temp_fou, (nz, ny, nx)) * (nx * ny * nz)   # unscaled C2R
        del temp_fou, tpad
        ccf = pad_zeros3d(ccf, trim_valid[0], trim_valid[1])          # padVal rows: negative = trim
        ccf = ccf / matlab_std(ccf)
        if mag is None and init_from_first:
            mag = ccf.copy()
            ang = np.ones(ccf.shape)
            second = np.full(ccf.shape, -np.inf)
        else:
            if mag is None:
                mag = np.zeros(ccf.shape)
                ang = np.zeros(ccf.shape)
                second = np.zeros(ccf.shape)
            rep = mag < ccf
            second = np.where(rep, mag, np.maximum(second, ccf))
            mag = np.where(rep, ccf, mag)
            ang = np.where(rep, aid, ang)
        aid += 1
    if return_margin:
        return mag, ang, mag - second
    return mag, ang


# --------------------------------------------------------------------------
# Batched Fourier-space alignment score and its gradient (this package's
# batched path; no reference counterpart: it replaces the per-orientation
# loop of BH_alignRaw3d_v2.m:1119-1419 for many particles). Restated here in
# FP64 on the FULL frequency grid with numpy's inverse FFT, independently of
# the device's half-grid pruned DFT.
# --------------------------------------------------------------------------


def batch_reference_spectrum(ref, table="c64"):
    """rfftn of the reference zero-padded to 2n with its centre n//2 at the
    origin; shape (2n, 2n, n+1) [z, y, x].

    table "f16": the device's half-precision table. Real and imaginary parts
    are scaled by 2^(14-e), where max(|Re|, |Im|) < 2^e (frexp), rounded to
    IEEE half (nearest even) and scaled back. The normalised score does not
    depend on the power-of-two scale, only on the rounding."""
    ref = np.asarray(ref, dtype=np.float64)
    n = ref.shape[0]
    m = 2 * n
    pad = np.zeros((m, m, m))
    idx = (np.arange(n) - n // 2) % m
    pad[np.ix_(idx, idx, idx)] = ref
    F = _rfftn(pad)
    if table == "f16":
        a = max(np.max(np.abs(F.real)), np.max(np.abs(F.imag)))
        e = int(np.frexp(a)[1]) if a > 0 else 0
        sc = 2.0 ** (14 - e)
        half = lambda x: (x * sc).astype(np.float16).astype(np.float64) / sc
        F = half(F.real) + 1j * half(F.imag)
    elif table != "c64":
        raise ValueError(table)
    return F


def _f2_texel(F2, n, tx, ty, tz):
    m = 2 * n
    neg = tx < 0
    x = np.where(neg, -tx, tx)
    y = np.where(neg, -ty, ty) % m
    z = np.where(neg, -tz, tz) % m
    v = F2[z, y, x]
    return np.where(neg, np.conj(v), v)


def batch_sample(F2, n, j):
    """Trilinear sample of the 2n spectrum at index coordinates j (3, ...),
    with the Hermitian mirror for negative x. Returns (f, grad[3])."""
    jx, jy, jz = j
    x0, y0, z0 = np.floor(jx).astype(int), np.floor(jy).astype(int), np.floor(jz).astype(int)
    fx, fy, fz = jx - x0, jy - y0, jz - z0
    c = {}
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                c[dz, dy, dx] = _f2_texel(F2, n, x0 + dx, y0 + dy, z0 + dz)
    lerp = lambda a, b, w: a + w * (b - a)
    c00 = lerp(c[0, 0, 0], c[0, 0, 1], fx)
    c01 = lerp(c[0, 1, 0], c[0, 1, 1], fx)
    c10 = lerp(c[1, 0, 0], c[1, 0, 1], fx)
    c11 = lerp(c[1, 1, 0], c[1, 1, 1], fx)
    c0, c1 = lerp(c00, c01, fy), lerp(c10, c11, fy)
    f = lerp(c0, c1, fz)
    gz = c1 - c0
    gy = lerp(c01 - c00, c11 - c10, fz)
    dxv = {(dz, dy): c[dz, dy, 1] - c[dz, dy, 0] for dz in (0, 1) for dy in (0, 1)}
    gx = lerp(lerp(dxv[0, 0], dxv[0, 1], fy), lerp(dxv[1, 0], dxv[1, 1], fy), fz)
    return f, (gx, gy, gz)


@dataclass
class BatchParticle:
    a: np.ndarray        # P_centred * bandpass, full grid FFT order, DC 0
    w: np.ndarray        # wedge_sym * bandpass (real), full grid
    A: float             # sum |a|^2
    kvec: np.ndarray     # (3, m) integer frequencies inside the support
    sel: tuple           # indices of those frequencies in the full grid


def batch_prepare(particle, bandpass, wedge, support_radius):
    p = np.asarray(particle, dtype=np.float64)
    n = p.shape[0]
    P = np.fft.fftn(np.fft.ifftshift(p))
    P.flat[0] = 0.0
    bp = np.asarray(bandpass, dtype=np.float64)
    if wedge is None:
        S = np.ones_like(bp)
    else:
        Wi = np.fft.ifftshift(np.asarray(wedge, dtype=np.float64))   # FFT order
        Wm = np.roll(Wi[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))        # W(-k)
        S = 0.5 * (Wi + Wm)
    f = np.fft.fftfreq(n) * n
    kz, ky, kx = np.meshgrid(f, f, f, indexing="ij")
    inside = (kx ** 2 + ky ** 2 + kz ** 2 <= support_radius ** 2)
    inside.flat[0] = False
    sel = np.nonzero(inside)
    kvec = np.stack([kx[sel], ky[sel], kz[sel]])
    a = P * bp
    # the particle norm runs over the whole grid (the bandpass sets its support)
    return BatchParticle(a, S * bp, float(np.sum(np.abs(a) ** 2)), kvec, sel)


def batch_corr_terms(F2, n, bp_: BatchParticle, R):
    """c(k) = F(2 R^T k) w(k) on the support, and a(k)."""
    M = 2.0 * np.asarray(R, dtype=np.float64).T
    j = M @ bp_.kvec
    f, g = batch_sample(F2, n, j)
    return f, g, M


def batch_score_map(F2, n, bp_: BatchParticle, R, peak_radius):
    """(score, shift) of one (particle, orientation): normalised CC over the
    candidate shifts |d| <= r, parabolic sub-pixel refinement (the device's
    estimator)."""
    f, _, _ = batch_corr_terms(F2, n, bp_, R)
    c = f * bp_.w[bp_.sel]
    X = np.zeros((n, n, n), dtype=np.complex128)
    X[bp_.sel] = bp_.a[bp_.sel] * np.conj(c)
    M = np.real(_ifftn(X)) * X.size                # sum_k Re[X e^{+2 pi i k.d/n}]
    D = float(np.sum(np.abs(c) ** 2))
    r = peak_radius
    best, bv = None, -np.inf
    for dz in range(-r, r + 1):
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                if dx * dx + dy * dy + dz * dz > r * r:
                    continue
                v = M[dz % n, dy % n, dx % n]
                if v > bv:
                    bv, best = v, (dx, dy, dz)
    sh = []
    for ax in range(3):
        d = list(best)
        if abs(d[ax]) >= r:
            sh.append(float(d[ax]))
            continue
        lo, hi = list(d), list(d)
        lo[ax] -= 1
        hi[ax] += 1
        a_ = M[lo[2] % n, lo[1] % n, lo[0] % n]
        c_ = M[hi[2] % n, hi[1] % n, hi[0] % n]
        den = a_ - 2 * bv + c_
        off = 0.0 if not den < 0 else min(0.5, max(-0.5, 0.5 * (a_ - c_) / den))
        sh.append(d[ax] + off)
    return bv / math.sqrt(bp_.A * D), np.array(sh)


def batch_eval(F2, n, bp_: BatchParticle, R, s, dR=None):
    """Score at (R, s) and, with dR (3 derivative matrices of R), the
    gradient [d/dparam_1..3, d/ds_x, d/ds_y, d/ds_z] (FP64, full grid)."""
    f, g, M = batch_corr_terms(F2, n, bp_, R)
    w = bp_.w[bp_.sel]
    a = bp_.a[bp_.sel]
    k = bp_.kvec
    e = np.exp(2j * np.pi * (k.T @ np.asarray(s, dtype=np.float64)) / n)
    Y = a * np.conj(f * w) * e
    num = float(np.sum(Y.real))
    D = float(np.sum(np.abs(f * w) ** 2))
    norm = math.sqrt(bp_.A * D)
    score = num / norm
    if dR is None:
        return score
    grad = np.zeros(6)
    for i in range(3):
        grad[3 + i] = float(np.sum((-Y.imag) * 2 * np.pi * k[i] / n)) / norm
    for q in range(3):
        dj = 2.0 * np.asarray(dR[q], dtype=np.float64).T @ k
        df = g[0] * dj[0] + g[1] * dj[1] + g[2] * dj[2]
        dnum = float(np.sum(np.real(a * np.conj(df * w) * e)))
        dD = float(np.sum(2.0 * np.real(np.conj(f * w) * df * w)))
        grad[q] = dnum / norm - num * dD / (2.0 * D * norm)
    return score, grad


# --------------------------------------------------------------------------
# Template-search pre-processing (alignment/BH_templateSearch3d_2.m:480-611)
# --------------------------------------------------------------------------


def gaussian_taps(F, sigma=3.0):
    """EMC_gaussianKernel([1, F], sigma) (masking/EMC_gaussianKernel.m: row
    vector on EMC_coordVectors origin 0, i.e. -F/2+0.5 .. F/2-0.5), sum 1."""
    v = np.arange(F) - (F - 1) / 2.0
    k = np.exp(-0.5 * (v / sigma) ** 2)
    return k / k.sum()


def matlab_conv_same(u, v):
    """MATLAB conv(u, v, 'same'): the full convolution from (1-based) index
    floor(numel(v)/2) + 1, numel(u) samples."""
    full = np.convolve(np.asarray(u, dtype=np.float64), np.asarray(v, dtype=np.float64))
    s = len(v) // 2
    return full[s:s + len(u)]


def _conv_same_axis(a, k, axis):
    """convn(a, k along `axis`, 'same') (zero outside a)."""
    F, n = len(k), a.shape[axis]
    s = F // 2
    ap = np.pad(a, [(F, F) if ax == axis else (0, 0) for ax in range(a.ndim)])
    out = np.zeros_like(a)
    for m in range(F):
        sl = [slice(None)] * a.ndim
        start = F + s - m                # out[i] = sum_m k[m] a[i + s - m]
        sl[axis] = slice(start, start + n)
        out += k[m] * ap[tuple(sl)]
    return out


def moving_average(img, F):
    """BH_movingAverage_2 (statistics/BH_movingAverage_2.m:1-25): padarray
    'symmetric' by ceil(F/2) (:10-11), EMC_convn with the separable kernel
    (testScripts/EMC_convn.m:54-56: 'same' along x, y, z), crop back with
    EMC_limits/EMC_resize (:18-19; equal left/right crops)."""
    pad = (F + 1) // 2
    a = np.pad(np.asarray(img, dtype=np.float64), pad, mode="symmetric")
    k = gaussian_taps(F)
    k = k / k.sum()                                                  # :14
    for axis in (2, 1, 0):                                           # x, y, z of an (nz,ny,nx) array
        a = _conv_same_axis(a, k, axis)
    return a[pad:-pad, pad:-pad, pad:-pad]


def local_normalize(img, F):
    """BH_templateSearch3d_2.m:558-561."""
    x = np.asarray(img, dtype=np.float64)
    avg = moving_average(x, F)
    avg2 = moving_average(x * x, F)
    with np.errstate(invalid="ignore"):
        return (x - avg) / np.sqrt(avg2 - avg ** 2)


def mask3d_rectangle(size, radius):
    """BH_mask3d('rectangle', size, radius, [0,0,0]) (masking/BH_mask3d.m:
    :102-108 window of ones 2r+13 per axis, :244-280 pad to size with the
    singleTaper and the even-size origin shift, :462-466 final singleTaper).
    size and radius are (x, y, z); returns (nz, ny, nx)."""
    size = np.asarray(size, dtype=np.int64)
    win_xyz = (2 * np.asarray(radius, dtype=np.float64) + 13).astype(np.int64)
    win = np.ones(win_xyz[::-1])
    pv = pad_val(win_xyz, size)
    shift = 1 - size % 2
    pre = pv[0] + shift
    post = pv[1] - shift
    pre_trim = np.where(pre < 0, -pre, 0)
    post_trim = np.where(post < 0, -post, 0)
    pre = np.where(pre >= 0, pre, 0)
    post = np.where(post >= 0, post, 0)
    nz, ny, nx = win.shape
    win = win[pre_trim[2]:nz - post_trim[2], pre_trim[1]:ny - post_trim[1], pre_trim[0]:nx - post_trim[0]]
    win = pad_zeros3d(win, pre, post, taper=True)
    return pad_zeros3d(win, [0, 0, 0], [0, 0, 0], taper=True)


def median_filter3(x, k):
    """medfilt3(x, [k k k]) with MATLAB's default 'symmetric' padding (the
    border sample repeated in the mirror, scipy.ndimage's 'reflect' mode)."""
    from scipy import ndimage
    return ndimage.median_filter(np.asarray(x), size=int(k), mode="reflect")


def tsearch_preprocess_chunk(chunk, bp_vals, pixel, stats_radius, sign, valid_mask=None, med_filt=0):
    """One chunk of BH_templateSearch3d_2.m:532-589: returns (out, tmp_sum).
    bp_vals = (HIGH_THRESH, HIGH_CUT, LOW_CUT); sign = -shouldBeCTF;
    med_filt 3/5/7: medfilt3 after the bandpass (:543-548)."""
    x = np.asarray(chunk, dtype=np.float64)
    x = x - x.mean()                                                  # :535
    x = x / np.sqrt(np.mean(x * x))                                   # :536 (rms)
    nz, ny, nx = x.shape
    bp = bandpass3d((nx, ny, nz), bp_vals[0], bp_vals[1], bp_vals[2], pixel)
    # fwdFFT(x,0,0,[bp,pixel]) then invFFT(.., 2) (fourierTransformer.m:92-158, :294-328)
    x = np.fft.irfftn(np.fft.rfftn(x) * bp[:, :, :nx // 2 + 1], s=x.shape, axes=(0, 1, 2))
    if med_filt:
        x = median_filter3(x, med_filt)                               # :543-548
    x = local_normalize(x, stats_radius)                              # :558-561
    m = np.ones_like(x) if valid_mask is None else np.asarray(valid_mask, dtype=np.float64)
    out = (sign * x) * m                                              # :586
    return out, float(np.sum(out[m > 0.1]))                           # :589


def tsearch_preprocess(tomogram, size_chunk, valid_area, n_iters, tomo_pre, tomo_post, valid_calc,
                       bp_vals, pixel, stats_radius, should_be_ctf=1, med_filt=0):
    """BH_templateSearch3d_2.m:397-398 (symmetric padding) and :480-611: the
    stack of pre-processed chunks (chunk corners in the padded tomogram) and
    the global variance. The reference accumulates the square of each chunk's
    masked SUM (:590), so globalVariance = sum(tmp_sum^2)/N - (sum(tmp_sum)/N)^2
    with N = chunks * numel(chunk); every chunk is divided by its square root."""
    t = np.asarray(tomogram, dtype=np.float64)
    p = [(int(tomo_pre[2]), int(tomo_post[2])), (int(tomo_pre[1]), int(tomo_post[1])),
         (int(tomo_pre[0]), int(tomo_post[0]))]
    t = np.pad(t, p, mode="symmetric")
    sc = [int(s) for s in size_chunk]
    va = [int(s) for s in valid_area]
    vmask = mask3d_rectangle(sc, np.asarray(valid_calc, dtype=np.float64) / 2.0)   # :402
    chunks, corners = [], []
    fx = fx2 = fn = 0.0
    for ix in range(int(n_iters[0])):
        for iy in range(int(n_iters[1])):
            for iz in range(int(n_iters[2])):
                cx, cy, cz = ix * va[0], iy * va[1], iz * va[2]
                sub = t[cz:cz + sc[2], cy:cy + sc[1], cx:cx + sc[0]]
                out, s = tsearch_preprocess_chunk(sub, bp_vals, pixel, stats_radius, -float(should_be_ctf), vmask,
                                                  med_filt)
                fx += s
                fx2 += s * s
                fn += float(np.prod(sc))
                chunks.append(out)
                corners.append((cx, cy, cz))
    gvar = fx2 / fn - (fx / fn) ** 2                                  # :603
    return [c / math.sqrt(gvar) for c in chunks], corners, gvar, vmask


# --------------------------------------------------------------------------
# Template-search peak extraction (alignment/BH_templateSearch3d_2.m:1036-1275)
# --------------------------------------------------------------------------


def gaussian3d_taps(size=5, sd=0.5):
    """1-D factor of BH_multi_gaussian3d(size, sd) (statistics/... :1-30):
    exp(-0.5 (x/sd)^2) on x = -floor(size/2).., normalised (the 3-D kernel is
    the outer product, normalised to sum 1)."""
    x = np.arange(size) - size // 2
    g = np.exp(-0.5 * (x / sd) ** 2)
    return g / g.sum()


def mask3d_sphere(size, radius):
    """BH_mask3d('sphere', size, radius, [0,0,0]) (masking/BH_mask3d.m:80-160,
    :462-466); size and radius (x, y, z); returns (nz, ny, nx)."""
    nx, ny, nz = (int(v) for v in size)
    rx, ry, rz = (float(v) for v in radius)
    z, y, x = np.meshgrid(np.arange(nz) - nz // 2, np.arange(ny) - ny // 2, np.arange(nx) - nx // 2,
                          indexing="ij")
    ell = lambda s: (x / (rx + s)) ** 2 + (y / (ry + s)) ** 2 + (z / (rz + s)) ** 2
    full = ell(0) <= 1
    w = full.astype(np.float64)
    taper = 0.5 + 0.5 * np.cos(np.arange(1, 7) * np.pi / 7)                 # :87
    for s in range(1, 7):                                                     # :145-151
        w = w + ((ell(s) <= 1).astype(np.float64) - w) * taper[s - 1]
    k = gaussian3d_taps(5, 0.5)
    for axis in (2, 1, 0):                                                    # convn 'same' (:152)
        w = _conv_same_axis(w, k, axis)
    w = w / w.max()                                                           # :153
    w[full] = 1.0                                                             # :155
    return pad_zeros3d(w, [0, 0, 0], [0, 0, 0], taper=True)                   # :462-466


def mask_shape_kernel_taps(size, kernel):
    """The default blurring kernel of EMC_maskShape (masking/EMC_maskShape.m:96-107):
    K = round((min(SIZE) kernel - 1) / 2) 2 + 1, at least 9; sigma keeps the
    Gaussian at 0.5 at ceil(K/2)/2 (EMC_gaussianKernel.m, origin 0, single
    precision), normalised to sum 1. Returns float32 taps."""
    K = int(matlab_round((min(size) * kernel - 1) / 2)) * 2 + 1
    K = max(K, 9)
    middle = math.ceil(K / 2) / 2
    sigma = np.float32(math.sqrt(-1 * middle ** 2 / (2 * math.log(0.5))))
    v = (np.arange(K, dtype=np.float32) - np.float32((K - 1) / 2))
    k = np.exp(np.float32(-0.5) * (v / sigma) ** 2).astype(np.float32)
    return (k / np.float32(k.sum(dtype=np.float32))).astype(np.float32)


def mask_shape(shape, size, radius, shift=(0.0, 0.0, 0.0), kernel=0.04):
    """EMC_maskShape(SHAPE, SIZE, RADIUS, METHOD, {'shift', shift; 'kernel',
    kernel}) for a 3-D SIZE (x, y, z), origin 1 and single precision
    (masking/EMC_maskShape.m:93-208): 'sphere' (:126-146), 'cylinder' (:148-185,
    the ellipse of x, y broadcast over 2 R_z + 1 slices about z = shift_z),
    'rectangle' (:187-198, |v| <= R). kernel False or 0: the binary shape;
    True: 0.04; a fraction: the Gaussian of mask_shape_kernel_taps, the shape's
    radius enlarged by ceil(K/2) (sphere, cylinder x/y) or floor(K/2)
    (rectangle, cylinder z), the outer ceil(K/2) layers of every face zeroed
    (EMC_resize 'force_taper' with a zero taper, :212), a separable 'same'
    convolution (EMC_convn), normalised by its maximum, values <= 0.001 set
    to 0 (:213-215). The coordinates are v = i - floor(n/2) - shift
    (EMC_coordVectors, origin 1). Returns (nz, ny, nx) float64."""
    f32 = np.float32
    nx, ny, nz = (int(v) for v in size)
    R = np.array(radius, dtype=np.float64).reshape(3)
    sh = np.array(shift, dtype=np.float64).reshape(3)
    if kernel is True:
        kernel = 0.04
    taps = mask_shape_kernel_taps(size, kernel) if kernel else None
    K = 0 if taps is None else len(taps)
    vx = (np.arange(nx) - nx // 2).astype(f32) - f32(sh[0])
    vy = (np.arange(ny) - ny // 2).astype(f32) - f32(sh[1])
    vz = (np.arange(nz) - nz // 2).astype(f32) - f32(sh[2])
    if shape == "sphere":
        R = R + math.ceil(K / 2)
        tx, ty, tz = ((vx / f32(R[0])) ** 2, (vy / f32(R[1])) ** 2, (vz / f32(R[2])) ** 2)
        m = ((tx[None, None, :] + ty[None, :, None]) + tz[:, None, None]) <= f32(1.0)
    elif shape == "cylinder":
        rxy = R[:2] + math.ceil(K / 2)
        rz = int(matlab_round(R[2])) + K // 2
        vx = (np.arange(nx) - nx // 2).astype(f32) - f32(sh[0])
        vy = (np.arange(ny) - ny // 2).astype(f32) - f32(sh[1])
        ell = ((vx / f32(rxy[0])) ** 2)[None, :] + ((vy / f32(rxy[1])) ** 2)[:, None] <= f32(1.0)
        iz = np.arange(nz) - nz // 2 - int(matlab_round(sh[2]))
        m = ell[None, :, :] & (np.abs(iz) <= rz)[:, None, None]
    elif shape == "rectangle":
        r = matlab_round(R).astype(int) + K // 2
        si = matlab_round(sh).astype(int)
        ix, iy, iz = (np.arange(n) - n // 2 - si[a] for a, n in enumerate((nx, ny, nz)))
        m = ((np.abs(ix) <= r[0])[None, None, :] & (np.abs(iy) <= r[1])[None, :, None]
             & (np.abs(iz) <= r[2])[:, None, None])
    else:
        raise ValueError(shape)
    out = m.astype(np.float64)
    if taps is not None:
        e = math.ceil(K / 2)
        out[:e], out[-e:] = 0.0, 0.0
        out[:, :e], out[:, -e:] = 0.0, 0.0
        out[:, :, :e], out[:, :, -e:] = 0.0, 0.0
        k = taps.astype(np.float64)
        for axis in (2, 1, 0):
            out = _conv_same_axis(out, k, axis)
        out = out / out.max()
        out[out <= 0.001] = 0.0
    return out


def tsearch_peaks(mag, ang, angle_list, removal_mask, rm_dim, mask_cutoff, high_thr, max_peaks, max_tries,
                  convention="Bah"):
    """The greedy loop of BH_templateSearch3d_2.m:1130-1272 for nPeaks = 1 on
    padded maps (nz, ny, nx); mag is copied. Returns (n, 7): x, y, z (1-based
    voxel + centre of mass - rm_dim), angles, score."""
    mag = np.array(mag, dtype=np.float64, copy=True)
    nz, ny, nx = mag.shape
    flat = mag.reshape(-1)          # x fastest, as MATLAB's column-major (x, y, z)
    peaks, tries = [], 0
    coord = int(np.argmax(flat))
    while len(peaks) < max_peaks and tries < max_tries and flat[coord] > high_thr:
        tries += 1
        z, r = divmod(coord, nx * ny)
        y, x = divmod(r, nx)
        a = int(ang[z, y, x])
        if a > 0:
            box = mag[z - 1:z + 2, y - 1:y + 2, x - 1:x + 2]
            d = np.arange(-1, 2)
            s = box.sum()
            com = [np.sum(box * d[None, None, :]) / s, np.sum(box * d[None, :, None]) / s,
                   np.sum(box * d[:, None, None]) / s]
            angs = np.asarray(angle_list[a - 1], dtype=np.float64)
            peaks.append([x + 1 + com[0] - rm_dim, y + 1 + com[1] - rm_dim, z + 1 + com[2] - rm_dim,
                          angs[0], angs[1], angs[2], flat[coord]])
            R = define_matrix(angs, convention, "forward")
            rot = xform3d(removal_mask, R, np.zeros(3), True)
            sl = (slice(z - rm_dim, z + rm_dim + 1), slice(y - rm_dim, y + rm_dim + 1),
                  slice(x - rm_dim, x + rm_dim + 1))
            mag[sl] = mag[sl] * (rot < mask_cutoff)
        else:
            flat[coord] = 0.0
        coord = int(np.argmax(flat))
    return np.array(peaks).reshape(-1, 7)


def tsearch_peaks_multi(mag, ang, angle_list, removal_mask, rm_dim, mask_cutoff, high_thr, max_peaks, max_tries,
                        n_peaks, sampling_rate, normals, convention="Bah"):
    """The loop of BH_templateSearch3d_2.m:1130-1272 with nPeaks extra
    columns (peakMat of :1097, rows of 10 nPeaks). Centre of mass over the
    box of half edge bDist = 1 + round(log(nPeaks)) (:1156-1180); the peaks
    2..nPeaks of :1198-1240, with MATLAB's indexing kept: magBox and angBox
    are linear over (x, y, z) (numpy's C order over (z, y, x)), Ang(cAng)
    indexes the whole padded id map with the box's index; randn(1) is
    normals[row][iPeak - 2][0..2]."""
    mag = np.array(mag, dtype=np.float64, copy=True)
    ang = np.asarray(ang)
    ang_flat = ang.reshape(-1)
    nz, ny, nx = mag.shape
    flat = mag.reshape(-1)
    b = 1 + int(np.floor(np.log(n_peaks) + 0.5))          # MATLAB round, log > 0 never ends on .5
    d = np.arange(-b, b + 1)
    rows, tries = [], 0
    coord = int(np.argmax(flat))
    while len(rows) < max_peaks and tries < max_tries and flat[coord] > high_thr:
        tries += 1
        z, r = divmod(coord, nx * ny)
        y, x = divmod(r, nx)
        if ang_flat[coord] > 0:
            mag_box = mag[z - b:z + b + 1, y - b:y + b + 1, x - b:x + b + 1].astype(np.float32)
            ang_box = ang[z - b:z + b + 1, y - b:y + b + 1, x - b:x + b + 1]
            mb = mag_box.astype(np.float64)
            s = mb.sum()
            cen = [x + 1 + np.sum(mb * d[None, None, :]) / s - rm_dim,
                   y + 1 + np.sum(mb * d[None, :, None]) / s - rm_dim,
                   z + 1 + np.sum(mb * d[:, None, None]) / s - rm_dim]
            row = np.zeros(10 * n_peaks)
            top = int(ang_flat[coord])
            row[0:3] = sampling_rate * np.asarray(cen)
            row[3:6] = angle_list[top - 1]
            row[9] = flat[coord]
            snr = 0.0
            if n_peaks > 1:
                possible = mag_box.reshape(-1).copy()
                abox = ang_box.reshape(-1)
                possible[abox == top] = 0
                n_random = 2
                for ip in range(2, n_peaks + 1):
                    use_random = False
                    if np.any(possible != 0):
                        c_ang = int(np.argmax(possible))
                        top = int(abox[c_ang])
                        snr = float(np.float32(np.mean(possible[abox == top], dtype=np.float64)))
                        possible[abox == top] = 0
                        if top <= 0 or ang_flat[c_ang] <= 0:
                            use_random = True
                            snr = np.mean(row[9:10 * (ip - 1):10]) * 0.75
                        else:
                            i_angles = np.asarray(angle_list[int(ang_flat[c_ang]) - 1], dtype=np.float64)
                    else:
                        use_random = True
                    if use_random:
                        i_angles = np.asarray(normals[len(rows)][ip - 2], dtype=np.float64) * n_random ** 2 + row[0:3]
                        if n_random < 10:
                            n_random += 1
                    k = 10 * (ip - 1)
                    row[k:k + 3] = row[0:3]
                    row[k + 3:k + 6] = i_angles
                    row[k + 9] = snr
            rows.append(row)
            R = define_matrix(np.asarray(angle_list[int(ang_flat[coord]) - 1], dtype=np.float64), convention,
                              "forward")
            rot = xform3d(removal_mask, R, np.zeros(3), True)
            sl = (slice(z - rm_dim, z + rm_dim + 1), slice(y - rm_dim, y + rm_dim + 1),
                  slice(x - rm_dim, x + rm_dim + 1))
            mag[sl] = mag[sl] * (rot < mask_cutoff)
        else:
            flat[coord] = 0.0
        coord = int(np.argmax(flat))
    return np.array(rows).reshape(-1, 10 * n_peaks)


# --------------------------------------------------------------------------
# (f)2 MRC files (@MRCImage/private/readHeader.m:43-186, getVolume.m:40-298,
# SAVE_IMG.m, private/writeHeader.m, private/setVolumeAndHeaderFromVolume.m)
# and subtomogram extraction (logicals/BH_isWindowValid.m:42-65,
# alignment/BH_alignRaw3d_v2.m:1036-1086)
# --------------------------------------------------------------------------

# readHeader.m's read order: (name, numpy type code, count); "skip" fields
# are the unnamed bytes it reads and drops
_MRC_ORDER = [
    ("nX", "i4", 1), ("nY", "i4", 1), ("nZ", "i4", 1), ("mode", "i4", 1),
    ("nXStart", "i4", 1), ("nYStart", "i4", 1), ("nZStart", "i4", 1),
    ("mX", "i4", 1), ("mY", "i4", 1), ("mZ", "i4", 1),
    ("cellDimensionX", "f4", 1), ("cellDimensionY", "f4", 1), ("cellDimensionZ", "f4", 1),
    ("cellAngleX", "f4", 1), ("cellAngleY", "f4", 1), ("cellAngleZ", "f4", 1),
    ("mapColumns", "i4", 1), ("mapRows", "i4", 1), ("mapSections", "i4", 1),
    ("minDensity", "f4", 1), ("maxDensity", "f4", 1), ("meanDensity", "f4", 1),
    ("spaceGroup", "i2", 1), ("nSymmetryBytes", "i2", 1), ("nBytesExtended", "i4", 1),
    ("creatorID", "i2", 1), ("skip", "u1", 30), ("nBytesPerSection", "i2", 1), ("serialEMType", "i2", 1),
    ("skip", "u1", 20), ("imodStamp", "i4", 1), ("imodFlags", "i4", 1),
    ("idtype", "i2", 1), ("lens", "i2", 1), ("ndl", "i2", 1), ("nd2", "i2", 1), ("vdl", "i2", 1),
    ("vd2", "i2", 1), ("tiltAngles", "f4", 6), ("xOrigin", "f4", 1), ("yOrigin", "f4", 1),
    ("zOrigin", "f4", 1), ("map", "u1", 4), ("machineStamp", "u1", 4), ("densityRMS", "f4", 1),
    ("nLabels", "i4", 1),
]
_MRC_GETVOLUME_TYPES = {0: "i1", 1: "i2", 2: "f4", 6: "u2"}   # getVolume.m:150-178


def mrc_read_header(path):
    """Field by field in readHeader.m's order; the byte order flips when a
    size reads negative or all three exceed 65535 (:54-78)."""
    raw = open(path, "rb").read(1024)
    first = np.frombuffer(raw[:12], "<i4")
    swap = bool(np.any(first < 0) or np.all(first > 65535))
    bo = ">" if swap else "<"
    h, pos = {}, 0
    for name, t, cnt in _MRC_ORDER:
        dt = np.dtype(bo + t) if t != "u1" else np.dtype("u1")
        v = np.frombuffer(raw[pos:pos + dt.itemsize * cnt], dt)
        pos += dt.itemsize * cnt
        if name == "skip":
            continue
        h[name] = bytes(v) if t == "u1" else (v.tolist() if cnt > 1 else v[0].item())
    nl = max(0, min(10, h["nLabels"]))
    h["labels"] = [raw[pos + 80 * i: pos + 80 * (i + 1)] for i in range(nl)]
    h["byteOrder"] = bo
    h["dataIndex"] = 1024 + max(h["nBytesExtended"], 0)
    return h


def mrc_get_volume(path, i_range=None, j_range=None, k_range=None):
    """getVolume(MRCImage(path), [i0 i1], [j0 j1], [k0 k1]) -> float64 (k, j, i)
    array; mode 0 reads as signed bytes (the 'int8' precision of :151-154)."""
    h = mrc_read_header(path)
    t = np.dtype(h["byteOrder"] + _MRC_GETVOLUME_TYPES[h["mode"]])
    nx, ny, nz = h["nX"], h["nY"], h["nZ"]
    data = np.fromfile(path, dtype=t, count=nx * ny * nz, offset=h["dataIndex"]).reshape(nz, ny, nx)
    i0, i1 = i_range or (1, nx)
    j0, j1 = j_range or (1, ny)
    k0, k1 = k_range or (1, nz)
    return data[k0 - 1:k1, j0 - 1:j1, i0 - 1:i1].astype(np.float64)


def _matlab_std_default(v):
    """MATLAB std(X) without a dimension: along the first non-singleton
    dimension (a scalar gives 0), N-1 normalisation. ``v`` in MATLAB order."""
    axes = [a for a in range(v.ndim) if v.shape[a] != 1]
    if not axes:
        return np.zeros_like(v)
    return np.std(v, axis=axes[0], ddof=1, keepdims=True)


def mrc_save(volume, path, pixel=None, origin=None):
    """SAVE_IMG(MRCImage(single(volume)), path[, pixel[, origin]]): header of
    setVolumeAndHeaderFromVolume.m, statistics of SAVE_IMG.m (min, max, mean,
    and rms = std(std(std(vol))), each std along the first non-singleton
    dimension, so 0 for a volume with a singleton axis), nLabels written
    as 1 (writeHeader.m), the origin's one-value form = -size/2 set before
    the cell is scaled by the pixel size. Field by field in the read order."""
    v = np.asarray(volume, dtype=np.float32)
    if v.ndim == 2:
        v = v[None]
    nz, ny, nx = v.shape
    vd = v.astype(np.float64)
    rms = vd.transpose(2, 1, 0)     # MATLAB (nx, ny, nz)
    for _ in range(3):
        rms = _matlab_std_default(rms)
    rms = float(rms.reshape(-1)[0])
    px = np.ones(3) if pixel is None else np.broadcast_to(np.asarray(pixel, float), (3,))
    if origin is None:
        org = [0.0, 0.0, 0.0]
    elif np.size(origin) == 1:
        org = [-nx / 2, -ny / 2, -nz / 2]
    else:
        org = list(np.asarray(origin, float))
    h = {"nX": nx, "nY": ny, "nZ": nz, "mode": 2, "nXStart": 0, "nYStart": 0, "nZStart": 0,
         "mX": nx, "mY": ny, "mZ": nz,
         "cellDimensionX": nx * px[0], "cellDimensionY": ny * px[1], "cellDimensionZ": nz * px[2],
         "cellAngleX": 90, "cellAngleY": 90, "cellAngleZ": 90,
         "mapColumns": 1, "mapRows": 2, "mapSections": 3,
         "minDensity": vd.min(), "maxDensity": vd.max(), "meanDensity": vd.mean(),
         "spaceGroup": 1, "nSymmetryBytes": 0, "nBytesExtended": 0, "creatorID": 0,
         "nBytesPerSection": 0, "serialEMType": 0, "imodStamp": 1146047817, "imodFlags": 0,
         "idtype": 0, "lens": 0, "ndl": 0, "nd2": 0, "vdl": 0, "vd2": 0, "tiltAngles": [0.0] * 6,
         "xOrigin": org[0], "yOrigin": org[1], "zOrigin": org[2],
         "map": b"MAP ", "machineStamp": bytes([68, 65, 0, 0]), "densityRMS": rms, "nLabels": 1}
    out = bytearray()
    for name, t, cnt in _MRC_ORDER:
        if name == "skip":
            out += bytes(cnt)
        elif t == "u1":
            out += h[name]
        else:
            out += np.asarray(h[name], dtype="<" + t).reshape(cnt).tobytes()
    out += bytes(1024 - len(out))
    with open(path, "wb") as f:
        f.write(bytes(out))
        f.write(np.ascontiguousarray(v, dtype="<f4").tobytes())


def window_valid(volume_size, window, mask_radius, center):
    """BH_isWindowValid.m:42-65 -> (indices 2x3 1-based or None for 'noUse',
    pads 2x3, shifts 3)."""
    vs = np.asarray(volume_size, float)
    ws = np.asarray(window, float)
    c = np.asarray(center, float)
    min_size = (np.max(mask_radius) + 6) * 2
    lo_c = np.ceil((ws - 1) / 2)
    hi_c = lo_c - np.mod(ws - 1, 2)
    wc = np.floor(c)
    LOW = wc - lo_c
    TOP = vs - (wc + hi_c)
    pad_lo = (np.abs(LOW) + 1) * (LOW <= 0)
    pad_hi = np.abs(TOP) * (TOP <= 0)
    ind = np.stack([LOW + pad_lo, wc + hi_c - pad_hi]).astype(np.int64)
    pads = np.stack([pad_lo, pad_hi]).astype(np.int64)
    avail = ws - pads[0] - pads[1]
    if np.any(avail - min_size < -2):
        return None, pads, c - wc
    return ind, pads, c - wc


def extract_subtomo(tomo, window, mask_radius, center, taper=True):
    """The per-particle cut of BH_alignRaw3d_v2.m:1036-1086 on a (z, y, x)
    tomogram: window check, crop, BH_padZeros3d(.., 'singleTaper') with zero
    fill (taper False: 'single', BH_average3d.m:1345). Returns (window
    (z, y, x) float64 or None, shift (3,))."""
    nz, ny, nx = tomo.shape
    ind, pads, shift = window_valid((nx, ny, nz), window, mask_radius, center)
    if ind is None:
        return None, shift
    cut = np.asarray(tomo, np.float64)[ind[0, 2] - 1:ind[1, 2], ind[0, 1] - 1:ind[1, 1], ind[0, 0] - 1:ind[1, 0]]
    return pad_zeros3d(cut, pads[0], pads[1], 0.0, taper=taper), shift


def calc_rg(stack_ny, r_coords):
    """BH_multi_loadOrBuild.m:203-210."""
    r = np.asarray(r_coords, float)
    oY = math.ceil((stack_ny + 1) / 2)
    nY = r[2] - r[1] + 1
    dY = math.floor(r[1] + nY / 2) - oY
    return np.array([[r[0], nY, r[3]], matlab_round(np.array([-r[4], dY, r[5]]))])


# --------------------------------------------------------------------------
# (f)3 alignment outputs: the per-tomogram rows of BH_alignRaw3d_v2.m:
# 1862-1903 and metaData/BH_rawAlignmentsApply.m:40-89
# --------------------------------------------------------------------------


def align_resume_text(rows, n_peaks=1):
    """alignResume/<prefix>/<tomo>.txt (:1894-1903): per row and peak
    '%d %d %6.3f %6.3f %6.3f %6.6f %6.6f %6.3f %6.3f %6.3f ' then a newline
    (MATLAB prints a non-integer under %d in %e form)."""
    def d(v):
        v = float(v)
        return "%d" % int(v) if v == int(v) else ("%e" % v)
    out = []
    for r in np.asarray(rows, float).reshape(len(rows), -1):
        line = ""
        for p in range(n_peaks):
            b = r[10 * p:10 * p + 10]
            line += "%s %s %6.3f %6.3f %6.3f %6.6f %6.6f %6.3f %6.3f %6.3f " % (
                d(b[0]), d(b[1]), b[2], b[3], b[4], b[5], b[6], b[7], b[8], b[9])
        out.append(line + "\n")
    return "".join(out)


def unique_particle_rows(sort_ccc):
    """:1866-1884: drop rows whose particle id (column 2) is 0, then keep the
    first row of every particle id in order ([~,a] = unique(.., 'stable'))."""
    s = np.asarray(sort_ccc, float)
    s = s[s[:, 1] != 0]
    seen, keep = set(), []
    for i, pid in enumerate(s[:, 1]):
        if pid not in seen:
            seen.add(pid)
            keep.append(i)
    return s[keep]


def raw_alignments_apply(geometry, best_angles, sampling, n_peaks, convention="Bah",
                         update_weights=False, update_class=False):
    """BH_rawAlignmentsApply.m:40-89, particle by particle: for every
    particle whose class (column 26) is not -9999, the row of its id (column
    4) in the tomogram's alignment rows; none -> class -9999; else rotation
    columns 17:25 <- reshape(old * define_matrix(angles, conv, 'inv')),
    position 11:13 += shifts * sampling, column 1 <- CCC, column 2 <- the
    weight column if asked, column 26 <- the reference id if asked."""
    out = {}
    for tomo, pos in geometry.items():
        pos = np.array(pos, dtype=float, copy=True)
        new = np.asarray(best_angles.get(tomo, np.zeros((0, 10 * n_peaks))), float)
        if new.size:
            for p in range(n_peaks):
                o, q = 26 * p, 10 * p
                for i in range(pos.shape[0]):
                    if pos[i, 25 + o] == -9999:
                        continue
                    hits = np.flatnonzero(new[:, 1] == pos[i, 3 + o])
                    if hits.size == 0:
                        pos[i, 25 + o] = -9999
                        continue
                    row = new[hits[0]]
                    old = pos[i, 16 + o:25 + o].reshape(3, 3, order="F")
                    rot = old @ define_matrix(row[2 + q:5 + q], convention, "inv")
                    pos[i, 16 + o:25 + o] = rot.reshape(-1, order="F")
                    pos[i, 10 + o:13 + o] += row[7 + q:10 + q] * sampling
                    pos[i, 0 + o] = row[5 + q]
                    if update_weights:
                        pos[i, 1 + o] = row[6 + q]
                    if update_class:
                        pos[i, 25 + o] = row[0 + q]
        out[tomo] = pos
    return out


# --------------------------------------------------------------------------
# (f)4 averaging: the per-particle body of transformations/BH_average3d.m:
# 1417-1441 and the half-set sums of :1465-1490
# --------------------------------------------------------------------------


def average_particles(windows, mats, shifts, halves, weights, interp_mask, wedges=None, symmetry="C1"):
    """Particle by particle in FP64: P = interpolator(window, M, shift, 'inv',
    symmetry); P = (P - mean(P(mask > 0.01))) / rms(..) .* mask; a particle
    whose normalised mean is not finite is skipped (:1445-1462); else added
    times its weight to the odd (half 1) or even (2) sum, and its rotated
    wedge likewise. Returns (odd, eve, wedge_odd, wedge_eve, used)."""
    mask = np.asarray(interp_mask, np.float64)
    binary = mask > 0.01
    odd = np.zeros(mask.shape)
    eve = np.zeros(mask.shape)
    wshape = None if wedges is None else np.asarray(wedges[0]).shape
    wodd = None if wedges is None else np.zeros(wshape)
    weve = None if wedges is None else np.zeros(wshape)
    used = []
    for i, win in enumerate(windows):
        P = interp3d(win, np.asarray(mats[i], float), shifts[i], "Bah", "inv", symmetry)
        P = P - P[binary].mean()
        with np.errstate(divide="ignore", invalid="ignore"):
            P = P / math.sqrt(float(np.mean(P[binary] ** 2)))
            P = P * mask
        ok = bool(np.isfinite(P.mean()))
        used.append(ok)
        if not ok:
            continue
        if halves[i] == 1:
            odd += P * weights[i]
        else:
            eve += P * weights[i]
        if wedges is not None:
            W = interp3d(wedges[i], np.asarray(mats[i], float), [0, 0, 0], "Bah", "inv", symmetry) * weights[i]
            if halves[i] == 1:
                wodd += W
            else:
                weve += W
    return odd, eve, wodd, weve, np.array(used)
