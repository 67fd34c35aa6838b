"""Pins of the FP64 oracle (CPU). Each test checks the oracle against a
closed form, an invariant, an independent library routine or a hand-worked
case, chosen so a plausible mistake (dropped term, wrong sign, swapped
order, transposed operand, off-by-one centre) fails."""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import emc_oracle as O
from tests.conftest import load_tlt


# ------------------------------------------------------------ define_matrix

@pytest.mark.parametrize("conv", ["Bah", "SPIDER", "Helical", "IMOD"])
def test_define_matrix_orthonormal(conv, rng):
    for _ in range(5):
        a = rng.uniform(-180, 180, 3)
        for d in ("forward", "inv"):
            R = O.define_matrix(a, conv, d)
            assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
            assert abs(np.linalg.det(R) - 1.0) < 1e-14


def test_define_matrix_hand_cases():
    # forward [90,0,0] = Rz(-90): x -> -y
    R = O.define_matrix([90, 0, 0], "Bah", "forward")
    assert np.allclose(R, [[0, 1, 0], [-1, 0, 0], [0, 0, 1]], atol=1e-15)
    # forward [90,90,0] = Rz(0) Rx(-90) Rz(-90); worked by hand. The swapped
    # product Rz(-90) Rx(-90) gives [[0,0,1],[-1,0,0],[0,-1,0]] instead.
    R = O.define_matrix([90, 90, 0], "Bah", "forward")
    assert np.allclose(R, [[0, 1, 0], [0, 0, 1], [1, 0, 0]], atol=1e-15)
    # TILT 'forwardVector' about Y
    a = math.radians(30)
    T = O.define_matrix([30], "TILT", "forwardVector")
    assert np.allclose(T, [[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])


def test_define_matrix_inv_is_forward_transpose(rng):
    for _ in range(10):
        a = rng.uniform(-180, 180, 3)
        assert np.allclose(O.define_matrix(a, "Bah", "inv"), O.define_matrix(a, "Bah", "forward").T,
                           atol=1e-15)


# ------------------------------------------------------------- eulerSearch

@pytest.mark.parametrize("args,count", [((180, 10, 180, 10), 15244), ((180, 15, 180, 15), 4600),
                                        ((180, 7, 180, 7), 43784), ((15, 3, 15, 3), 1045),
                                        ((15, 1, 15, 1), 23281), ((180, 10, 10, 10), 1236)])
def test_euler_search_counts(args, count):
    g = O.euler_search("C1", *args)
    assert int(g.n_at_theta.sum()) == count
    assert g.orientations().shape == (count, 3)


def test_euler_search_small_hand_case():
    g = O.euler_search("C1", 0, 0, 10, 5)
    # theta=0 only, one phi (0), psi = -10:5:10
    assert np.allclose(g.orientations(), [[0, 0, p] for p in (-10, -5, 0, 5, 10)])
    g = O.euler_search("C4", 30, 15, 180, 30)
    # C4 caps psi to +-45: psi = -45:30:45 -> -45,-15,15,45
    assert np.allclose(g.psi, [-45, -15, 15, 45])
    # theta 15: phi_step = 15/sin(15) = 57.96 -> 360/floor(6.21+0.5) = 60 -> 6 phis
    assert len(g.phi[1]) == 6 and np.isclose(g.phi[1][1], 60.0)


# ------------------------------------------------------------------ xform3d

def test_xform_identity_and_integer_shift(rng):
    v = rng.standard_normal((6, 7, 8))
    assert np.array_equal(O.xform3d(v, np.eye(3), [0, 0, 0], True), v)
    # forward: p = u - t, so out(x) = in(x - 1) along x
    s = O.xform3d(v, np.eye(3), [1, 0, 0], True)
    assert np.array_equal(s[:, :, 1:], v[:, :, :-1]) and not s[:, :, 0].any()
    # inverse: p = u + t, so out(z) = in(z + 2)
    s = O.xform3d(v, np.eye(3), [0, 0, 2], False)
    assert np.array_equal(s[:-2], v[2:]) and not s[-2:].any()


@pytest.mark.parametrize("n", [7, 8])
def test_xform_90deg_permutations(n, rng):
    v = rng.standard_normal((n, n, n))
    c = n // 2
    z, y, x = np.meshgrid(range(n), range(n), range(n), indexing="ij")
    # rows of M act on u = index - n//2; sample at M u + c
    cases = {
        "z": (np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]]), (c - (y - c), c + (x - c), z)),
        "x": (np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0.0]]), (x, c - (z - c), c + (y - c))),
        "y": (np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0.0]]), (c + (z - c), y, c - (x - c))),
    }
    for M, (px, py, pz) in cases.values():
        out = O.xform3d(v, M, [0, 0, 0], True)
        ok = (px >= 0) & (px < n) & (py >= 0) & (py < n) & (pz >= 0) & (pz < n)
        exp = np.zeros_like(v)
        exp[ok] = v[pz[ok], py[ok], px[ok]]
        assert np.array_equal(out, exp)


def test_xform_matches_scipy_linear(rng):
    """Independent routine: scipy map_coordinates order 1, 'grid-constant'
    (samples outside the grid are 0 and are interpolated, texture border)."""
    v = rng.standard_normal((5, 6, 7))
    M = O.define_matrix([23, 47, -71], "Bah", "forward")
    t = np.array([0.3, -0.7, 1.1])
    out = O.xform3d(v, M, t, True)
    nz, ny, nx = v.shape
    z, y, x = np.meshgrid(range(nz), range(ny), range(nx), indexing="ij")
    u = np.stack([x - nx // 2, y - ny // 2, z - nz // 2]).reshape(3, -1) - t[:, None]
    p = M @ u
    coords = np.stack([p[2] + nz // 2, p[1] + ny // 2, p[0] + nx // 2])
    ref = ndimage.map_coordinates(v, coords, order=1, mode="grid-constant", cval=0.0).reshape(v.shape)
    assert np.allclose(out, ref, atol=1e-12)


def test_xform_fwd_inv_sign_identity(rng):
    """fwd (M, t) samples the same points as inv (M, -M t)."""
    v = rng.standard_normal((8, 8, 8))
    M = O.define_matrix([10, 70, 200], "Bah", "inv")
    t = np.array([1.25, -0.5, 2.0])
    assert np.allclose(O.xform3d(v, M, t, True), O.xform3d(v, M, -M @ t, False), atol=1e-12)


def test_interp3d_reads_a_geometry_vector_column_major(rng):
    v = rng.standard_normal((8, 9, 10))
    M = O.define_matrix([20, 35, -60], "Bah", "inv")
    a = O.interp3d(v, M.reshape(-1, order="F"), [0.5, 0, -1], "Bah", "inv")
    assert np.array_equal(a, O.interp3d(v, M, [0.5, 0, -1], "Bah", "inv"))
    assert np.array_equal(a, O.interp3d(v, [20, 35, -60], [0.5, 0, -1], "Bah", "inv"))


def test_interp3d_symmetry_average_invariant(rng):
    v = rng.standard_normal((9, 9, 9))
    M = O.define_matrix([15, 30, 45], "Bah", "forward")
    S = O.symmetry_matrices("C4")
    a = O.interp3d(v, M, [0, 0, 0], "Bah", "forward", "C4")
    b = O.interp3d(v, M @ S[1], [0, 0, 0], "Bah", "forward", "C4")
    assert np.allclose(a, b, atol=1e-10)


# ------------------------------------------------------------------- CTF

def test_ctf_zero_at_analytic_frequency():
    lam, df, nx = 0.019687, 30000.0, 64
    k1 = math.sqrt(1.0 / (lam * df))          # sin(-pi lam df k^2) = 0 at k^2 = 1/(lam df)
    pix = 5.0 / (nx * k1)                     # puts k1 on the 5th Fourier pixel
    p = O.CtfParams(pix, lam, 0.0, 0.0, df, df, 0.0)
    img = O.ctf2d(nx, nx, p, half=False, sq=False, centered=True)
    assert abs(img[nx // 2, nx // 2 + 5]) < 1e-9 and abs(img[nx // 2 + 5, nx // 2]) < 1e-9
    assert abs(img[nx // 2, nx // 2 + 3]) > 0.1
    # DC: sin(-amp_phase) = 0 for amp 0
    assert img[nx // 2, nx // 2] == 0.0


def test_ctf_astigmatism_convention_and_weights():
    base = dict(pixel=2.0, wavelength=0.019687, cs_mm=2.7, amp=0.1, df1=32000.0, df2=28000.0)
    a = O.ctf2d(32, 32, O.CtfParams(astig_deg=200.0, **base), sq=True)
    b = O.ctf2d(32, 32, O.CtfParams(astig_deg=20.0, **base), sq=True)
    assert np.allclose(a, b, atol=1e-12)       # -90..90 wrap of the angle
    w = O.ctf2d(32, 32, O.CtfParams(astig_deg=20.0, **base), sq=True, weighted=True,
                occupancy=0.7, exposure=0.0)
    x = np.arange(32) - 16
    rw = 0.7 * np.where(x == 0, 0.2, np.abs(x))
    assert np.allclose(w, b * rw[None, :], atol=1e-12)
    # exposure filter at one point: exp(-E/2 / (0.24499 k^-1.665 + 2.8141))
    e = O.ctf2d(32, 32, O.CtfParams(astig_deg=20.0, **base), sq=True, weighted=True,
                occupancy=1.0, exposure=40.0)
    k = 5 / (2.0 * 32)
    f = math.exp(-20.0 / (0.24499 * k ** -1.665 + 2.8141))
    assert math.isclose(e[16, 21], b[16, 21] * 5 * f, rel_tol=1e-9)


def _tilt(t, df=30000.0):
    return dict(pixel=2.5, wavelength=0.019687, cs=2.7, df1=df, df2=df, astig=0.0, amp=0.1, tilt=t,
                exposure=0.0, occupancy=1.0)


def test_sf3d_single_zero_tilt_is_slab_times_ctf():
    n = 16
    w = O.sf3d((n, n, n), [_tilt(0.0)])
    img = O.ctf2d(n, n, O.CtfParams(2.5, 0.019687, 2.7, 0.1, 30000.0, 30000.0, 0.0), sq=True,
                  weighted=True, occupancy=1.0, exposure=0.0, spacing_from_full_nx=False)
    z = np.arange(n) - n // 2 + 0.5          # even nz: tw = w + 1/2
    zw = np.where(np.abs(z) <= 4, 0.25 * (0.5 + 0.5 * np.cos(np.pi * z / 4)), 0.0)
    assert np.allclose(w, zw[:, None, None] * img[None, :, :], atol=1e-12)


def test_sf3d_missing_wedge_is_empty():
    n = 32
    tilts = [_tilt(t) for t in np.arange(-60, 61, 3.0)]
    w = O.sf3d((n, n, n), tilts)
    # along kz at kx = 0 nothing is sampled beyond the slab reach of a 60 deg
    # tilt: |0.5 + w cos60| > 4  <=>  w >= 8 or w <= -10
    col = w[:, :, n // 2]
    assert not col[n // 2 + 8:].any() and not col[: n // 2 - 9].any()
    assert col[n // 2].any()


def test_weight_mask_on_reference_tlt_fixture():
    T = load_tlt("riboNew11_ali1_ctf.tlt")
    w = O.weight_mask_mex((16, 16, 16), 4, T, [100.0, 120.0, 40.0], [[400, 400, 200], [0, 0, 0]])
    assert np.isfinite(w).all() and abs(w.min() - 1e-4) < 1e-15
    tilts = O.tlt_to_tilts(T, 4, [100.0, 120.0, 40.0], [[400, 400, 200], [0, 0, 0]], (16, 16, 16))
    # the particle sits above the tilt axis: defocus changes with the tilt
    # by -(x sin a ... ) ; at the 0-deg-like tilt it equals the mean defocus
    # shifted by z*pixel (hand form of BH_weightMaskMex.m:33-37)
    i0 = int(np.argmin([abs(t["tilt"]) for t in tilts]))
    row = T[np.argsort(T[:, 0], kind="stable")][i0]
    pix = row[15] * 4 * 1e10
    prj = np.array([100.0, 120.0, 40.0]) - np.ceil((np.array([400, 400, 200]) + 1) / 2)
    a = math.radians(row[3])
    zc = -math.sin(a) * prj[0] + math.cos(a) * prj[2]
    assert math.isclose(tilts[i0]["df1"], -row[14] * 1e10 - zc * pix + row[11] * 1e10, rel_tol=1e-12)


# --------------------------------------------------------------- bandpass

def test_bandpass_closed_form():
    n = 32
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5)
    # DC equals HIGH_THRESH: highRoll^2 = -hc^2 / (2 ln thr)
    assert math.isclose(bp[0, 0, 0], 1e-3, rel_tol=1e-12)
    lc, hc = 2.5 / 8.0, 2.5 / 40.0
    r = O.fft_radius((n, n, n))
    inside = (r > hc) & (r < lc)
    assert np.all(bp[inside] == 1.0)
    # one lowRoll beyond lowCut: exp(-1/2)
    lroll = 2.0 / n
    sel = np.isclose(r, lc + lroll)
    if sel.any():
        assert np.allclose(bp[sel], math.exp(-0.5))
    # a voxel on the x axis at radius 12/32
    assert math.isclose(bp[0, 0, 12], math.exp(-((12 / 32 - lc) ** 2) / (2 * lroll ** 2)), rel_tol=1e-12)


# ----------------------------------------------- band-limit / normalise / pad

def test_fwd_fft_options_closed_forms(rng):
    """fourierTransformer.fwdFFT options: doNorm 1/2 scale by N^-1/2 and N^-1
    (MATLAB's fftn is the unnormalised DFT, pinned on a delta); centring on a
    volume whose x=0 and Nyquist planes hold only the DC term (an offset plus
    a cosine along x) leaves F(0) = 0 and, by Parseval, a full-grid
    sum |F|^2 of exactly 1; centring ignores doNorm."""
    nz, ny, nx = 6, 10, 16
    d = np.zeros((nz, ny, nx))
    d[0, 0, 0] = 1.0
    assert np.allclose(O.fwd_fft(d), 1.0)
    n = d.size
    v = rng.standard_normal((nz, ny, nx))
    assert np.allclose(O.fwd_fft(v, 1), np.fft.rfftn(v) / math.sqrt(n), rtol=0, atol=1e-12)
    assert np.allclose(O.fwd_fft(v, 2), np.fft.rfftn(v) / n, rtol=0, atol=1e-12)
    g = rng.standard_normal((nz, ny, 1))
    x = np.arange(nx)
    vol = 3.0 + g * np.cos(2 * np.pi * 3 * x / nx)
    for norm in (0, 1, 2):
        ft = O.fwd_fft(vol, norm, True)
        assert ft.flat[0] == 0
        full = np.sum(np.abs(ft[..., 0]) ** 2) + 2 * np.sum(np.abs(ft[..., 1:nx // 2]) ** 2) + \
            np.sum(np.abs(ft[..., nx // 2]) ** 2)
        assert math.isclose(full, 1.0, rel_tol=1e-12)
    # half bandpass multiplies plane-wise; ctf multiplies before the norm
    bp = rng.uniform(0, 1, (nz, ny, nx // 2 + 1))
    assert np.allclose(O.fwd_fft(v, 1, False, bp, 2 * bp), 2 * bp * bp * np.fft.rfftn(v) / math.sqrt(n))


def test_blcn_invariants(rng):
    img = rng.standard_normal((8, 8, 8)) + 3.0
    F = O.band_limit_center_normalize(img, None, [[0, 0, 0], [0, 0, 0]])
    assert F.flat[0] == 0
    assert math.isclose(np.sum(np.abs(F) ** 2), F.size ** 2, rel_tol=1e-12)
    # padding with the mean: the back-transform is (padded - mean) up to scale
    img = rng.standard_normal((6, 6, 6))
    F = O.band_limit_center_normalize(img, None, [[1, 1, 1], [1, 1, 1]])
    back = np.real(np.fft.ifftn(F))
    padded = np.full((8, 8, 8), img.mean())
    padded[1:7, 1:7, 1:7] = img
    d = padded - padded.mean()
    assert np.corrcoef(back.ravel(), d.ravel())[0, 1] > 1 - 1e-12


def test_pad_zeros3d_trim_and_taper(rng):
    img = rng.standard_normal((10, 10, 10))
    p = O.pad_zeros3d(img, [2, 1, 0], [0, 3, 1], extrap_val=0.5)
    assert p.shape == (11, 14, 12)
    assert np.array_equal(p[0:10, 1:11, 2:12], img) and np.all(p[10] == 0.5)
    t = O.pad_zeros3d(img, [-1, -1, -1], [-1, -1, -1])
    assert np.array_equal(t, img[1:-1, 1:-1, 1:-1])
    ones = np.ones((16, 16, 16))
    tp = O.pad_zeros3d(ones, [0, 0, 0], [0, 0, 0], taper=True)
    e = 0.5 + 0.5 * math.cos(6 * math.pi / 7)
    assert math.isclose(tp[8, 8, 0], e) and math.isclose(tp[0, 0, 0], e ** 3)
    assert tp[8, 8, 8] == 1.0


# ----------------------------------------------------------- translational

def test_xcf_translational_integer_shift(rng):
    n = 16
    x = rng.standard_normal((n, n, n))
    s = (2, -3, 1)                                  # (x, y, z)
    y = np.roll(x, (s[2], s[1], s[0]), axis=(0, 1, 2))
    p = O.xcf_translational(np.fft.fftn(y), np.conj(np.fft.fftn(x)), None, (0, 0, 0))
    assert np.array_equal(p, np.array(s, dtype=float))


def test_xcf_translational_com_half_pixel():
    n = 24
    ax = np.fft.fftfreq(n) * n                       # circular coordinates
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    g = np.exp(-((x - 1.5) ** 2 + (y + 2.0) ** 2 + z ** 2) / (2 * 0.8 ** 2))
    p = O.xcf_translational(np.fft.fftn(g), np.ones((n, n, n)), None, (3, 3, 3))
    assert np.allclose(p, [1.5, -2.0, 0.0], atol=0.02)


def test_xcf_translational_mask_and_empty():
    n = 16
    m = np.zeros((n, n, n))
    m[0, 0, 3] = 5.0                                 # strong peak at +3 x
    m[0, 0, n - 2] = 1.0                             # weak peak at -2 x
    mask = np.ones((n, n, n))
    mask[n // 2, n // 2, n // 2 + 3] = 0.0           # mask out the strong one
    p = O.xcf_translational(np.fft.fftn(m), np.ones((n, n, n)), mask, (0, 0, 0))
    assert np.allclose(p, [-2, 0, 0])
    z = O.xcf_translational(np.zeros((n, n, n)), np.zeros((n, n, n)), None, (3, 3, 3))
    assert np.allclose(z, -(n // 2 + 1))             # [0,0,0] - ceil((N+1)/2) (:95-97)


def test_xcf_translational_com_box_error():
    n = 16
    m = np.zeros((n, n, n))
    m[n // 2 - 1, 0, 0] = 1.0                        # fftshifted to the map edge
    with pytest.raises(IndexError):
        O.xcf_translational(np.fft.fftn(m), np.ones((n, n, n)), None, (3, 3, 3))


# ------------------------------------------------------------- rotational

def test_xcf_rotational_invariants(rng):
    P = np.fft.fftn(rng.standard_normal((8, 8, 8)))
    assert math.isclose(O.xcf_rotational(P, np.conj(P), np.ones(P.shape), 1.0), 1.0, rel_tol=1e-12)
    Q = np.fft.fftn(rng.standard_normal((8, 8, 8)))
    w = rng.uniform(0, 1, P.shape)
    c = O.xcf_rotational(P, np.conj(Q), w, 1.0)
    assert abs(c) < 1
    assert math.isclose(O.xcf_rotational(3 * P, np.conj(Q), w, 1.0), c, rel_tol=1e-12)
    wr = np.full(P.shape, np.nan)
    assert math.isclose(O.xcf_rotational(P, np.conj(Q), w, wr), c, rel_tol=1e-12)


def _radial_shells(n, edges):
    """1-based linear indices of the voxels with edges[i] <= r < edges[i+1]
    (radius from the FFT origin of an (n, n, n) grid)."""
    f = np.fft.fftfreq(n) * n
    z, y, x = np.meshgrid(f, f, f, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z).ravel()
    return [np.flatnonzero((lo <= r) & (r < hi)) + 1 for lo, hi in zip(edges[:-1], edges[1:])]


def test_xcf_rotational_shells_closed_forms(rng):
    """Shell-weighted CCC (BH_multi_xcf_Rotational.m:50-70) against values
    the formula fixes: per-shell scale invariance (each shell's own
    denominator), f(1) = (1/(1+w^-1.5))^3, f(-1) with and without abs, a
    one-voxel shell at MATLAB index 1 and one full shell = f(plain CC)."""
    n = 8
    R = np.fft.fftn(rng.standard_normal((n, n, n)))
    W = rng.uniform(0.2, 1, R.shape)
    sh = _radial_shells(n, [0, 1.5, 3, 7])
    # P a different positive multiple of conj(R) on each shell: every c = 1
    P = np.conj(R).copy()
    for s, t in zip(sh, (0.5, 3.0, 20.0)):
        P.flat[s - 1] *= t
    w = 4.0
    f1 = (1.0 / (1.0 + 0.125)) ** 3
    assert math.isclose(O.xcf_rotational(P, R, W, W, [0, w] + sh), 3 * f1, rel_tol=1e-12)
    # the same P with one shared denominator would not give c = 1 per shell
    assert O.xcf_rotational(P, R, W, W) < 0.99
    # c = -1 on shell 2 only: f(-1) = (1/(-1+w^-1.5))^3 < 0; abs flips it
    Pm = P.copy()
    Pm.flat[sh[1] - 1] *= -1
    fm1 = (1.0 / (-1.0 + 0.125)) ** 3
    assert math.isclose(O.xcf_rotational(Pm, R, W, W, [0, w] + sh), 2 * f1 + fm1, rel_tol=1e-12)
    assert math.isclose(O.xcf_rotational(Pm, R, W, W, [1, w] + sh), 2 * f1 - fm1, rel_tol=1e-12)
    # one-voxel shell at MATLAB index 1 = flat element 0: c = cos(phase)
    Q = np.fft.fftn(rng.standard_normal((n, n, n)))
    Q.flat[0] = np.conj(R.flat[0]) * np.exp(0.3j)
    c = math.cos(0.3)
    assert math.isclose(O.xcf_rotational(Q, R, W, W, [0, w, [1]]), (c * c / (c + 0.125)) ** 3, rel_tol=1e-12)
    # one shell holding every voxel: f of the plain CC
    cc = O.xcf_rotational(Q, R, W, 1.0)
    full = np.arange(1, n ** 3 + 1)
    assert math.isclose(O.xcf_rotational(Q, R, W, 1.0, [0, w, full]), (cc * cc / (cc + 0.125)) ** 3,
                        rel_tol=1e-12)


# ------------------------------------------------------------- template search

def test_swap_quadrants_is_shift_and_sign():
    """swapQuadrants = circular shift by n/2+1 and (-1)^(n/2+1) per axis."""
    n = 8
    sq = O.swap_quadrants_half((n, n, n))
    v = np.random.default_rng(3).standard_normal((n, n, n))
    a = np.fft.irfftn(sq * np.fft.rfftn(v), s=(n, n, n), axes=(0, 1, 2))
    sh = n // 2 + 1
    b = np.roll(v, (sh, sh, sh), axis=(0, 1, 2)) * (-1) ** (3 * (n // 2 + 1))
    assert np.allclose(a, b, atol=1e-12)


def test_template_search_finds_inserted_template():
    rng = np.random.default_rng(7)
    # compact asymmetric template: every 90-degree rotation (plus the [1,1,1]
    # shift) stays inside the 8^3 box, so no orientation loses energy
    ax = np.arange(8) - 4
    zz, yy, xx = np.meshgrid(ax, ax, ax, indexing="ij")
    t = np.where(xx ** 2 + yy ** 2 + zz ** 2 <= 4, rng.uniform(0.2, 1.0, (8, 8, 8)), 0.0)
    grid = O.euler_search("C1", 90, 90, 90, 90).orientations()[:6]
    k = 4
    phi, theta, psi = grid[k]
    rot = O.interp3d(t, [phi, theta, psi - phi], [1, 1, 1], "Bah", "forward")
    chunk = rng.standard_normal((24, 24, 24)) * 0.01
    c = np.array([12, 11, 13])                       # (x, y, z)
    chunk[c[2] - 4:c[2] + 4, c[1] - 4:c[1] + 4, c[0] - 4:c[0] + 4] += rot
    pad = O.pad_val((8, 8, 8), (24, 24, 24))
    valid = O.pad_val((24, 24, 24), (16, 16, 16))
    # the reference negates the tomogram in preprocessing (:586); for chunk
    # sizes = 0 mod 4 swapQuadrants carries (-1)^3 which cancels it
    mag, ang = O.template_search_chunk(-chunk, t, grid, pad[0:2], valid[0:2])
    i = np.unravel_index(np.argmax(mag), mag.shape)
    assert ang[i] == k + 1
    # the peak lands on the centre of the inserted copy: its box centre c
    # plus the [1,1,1] forward shift of the rotated template; valid area at 4
    assert np.array_equal(np.array(i)[::-1] + 4, c + 1)


# ------------------------------------------- batched Fourier-space scoring

def _batched_case(n=24):
    from emclarity_amd import synthetic as S
    c = S.make_align_case(4, n=n, euler=(40.0, 30.0, 20.0), snr=None, n_blobs=10, blob_radius=0.22 * n,
                          ctf=False, max_tilt=None)
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5)
    F2 = O.batch_reference_spectrum(c.ref)
    return c, bp, F2


def test_batch_spectrum_fold_and_rotation_identity():
    c, bp, F2 = _batched_case()
    n = c.ref.shape[0]
    # at even index coordinates the 2x spectrum equals the unpadded spectrum
    # of the centred reference (zero padding = sinc interpolation)
    direct = np.fft.fftn(np.fft.ifftshift(c.ref.astype(np.float64)))
    k = np.array([[3, -2, 5], [0, 4, -7], [-6, 1, 2]]).T
    f, _ = O.batch_sample(F2, n, 2.0 * k)
    assert np.allclose(f, direct[k[2] % n, k[1] % n, k[0] % n], rtol=1e-10, atol=1e-9)
    # Hermitian fold: F(-j) = conj F(j) at a non-integer point
    j = np.array([[1.3], [-2.7], [0.4]])
    fa, _ = O.batch_sample(F2, n, j)
    fb, _ = O.batch_sample(F2, n, -j)
    assert np.allclose(fa, np.conj(fb), rtol=1e-12)


def test_batch_score_prefers_truth_and_gradient_matches_fd():
    c, bp, F2 = _batched_case()
    n = c.ref.shape[0]
    bpp = O.batch_prepare(c.particle, bp, None, 9.0)
    Rt = c.rotmat
    st, sht = O.batch_score_map(F2, n, bpp, Rt, 4)
    so, _ = O.batch_score_map(F2, n, bpp, O.define_matrix([0, 0, 0], "Bah", "inv"), 4)
    assert st > so and st > 0.95
    assert np.max(np.abs(sht - c.shift)) < 0.25
    # analytic gradient vs central differences (FP64): rotation increments
    # about x, y, z and the shifts
    G = [np.array([[0, 0, 0], [0, 0, -1], [0, 1, 0.0]]), np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0.0]]),
         np.array([[0, -1, 0], [1, 0, 0], [0, 0, 0.0]])]
    R0 = Rt @ O.define_matrix([5, 3, 2], "Bah", "forward")
    s0 = c.shift + np.array([0.3, -0.2, 0.1])
    _, g = O.batch_eval(F2, n, bpp, R0, s0, [R0 @ Gi for Gi in G])
    h = 1e-5
    for i in range(3):
        w = np.zeros(3)
        w[i] = h
        Rp = R0 @ (np.eye(3) + np.sin(h) * G[i] + (1 - np.cos(h)) * G[i] @ G[i])
        Rm = R0 @ (np.eye(3) - np.sin(h) * G[i] + (1 - np.cos(h)) * G[i] @ G[i])
        fd = (O.batch_eval(F2, n, bpp, Rp, s0) - O.batch_eval(F2, n, bpp, Rm, s0)) / (2 * h)
        assert abs(fd - g[i]) < 1e-4 * max(abs(g[i]), 1e-3)
    for i in range(3):
        e = np.zeros(3)
        e[i] = 1e-5
        fd = (O.batch_eval(F2, n, bpp, R0, s0 + e) - O.batch_eval(F2, n, bpp, R0, s0 - e)) / 2e-5
        assert abs(fd - g[3 + i]) < 1e-5 * max(abs(g[3 + i]), 1e-3)


def test_batch_f16_table_rounding_and_scale_invariance():
    """The half-precision table: every part within half an fp16 ulp of the
    FP64 spectrum (relative 2^-11, or the subnormal spacing), the score
    unchanged by a power-of-two scale of the table, and within 1e-4 of the
    complex64 table's score."""
    c, bp, F2 = _batched_case()
    n = c.ref.shape[0]
    H = O.batch_reference_spectrum(c.ref, "f16")
    a = max(np.max(np.abs(F2.real)), np.max(np.abs(F2.imag)))
    sc = 2.0 ** (14 - int(np.frexp(a)[1]))
    for x, h in ((F2.real, H.real), (F2.imag, H.imag)):
        assert np.all(np.abs(h - x) <= np.maximum(2.0 ** -11 * np.abs(x), 2.0 ** -25 / sc) * (1 + 1e-12))
        assert np.max(np.abs(x)) * sc < 2.0 ** 14
    assert not np.array_equal(H, F2)   # rounding happened
    bpp = O.batch_prepare(c.particle, bp, None, 9.0)
    R = c.rotmat @ O.define_matrix([3, 2, 1], "Bah", "forward")
    s64, sh64 = O.batch_score_map(F2, n, bpp, R, 4)
    s16, sh16 = O.batch_score_map(H, n, bpp, R, 4)
    s16x, sh16x = O.batch_score_map(8.0 * H, n, bpp, R, 4)
    assert abs(s16 - s16x) < 1e-12 and np.allclose(sh16, sh16x, atol=1e-12)
    assert abs(s16 - s64) < 1e-4 * abs(s64) and np.max(np.abs(sh16 - sh64)) < 1e-3


# ------------------------------------------ template-search pre-processing

def test_matlab_conv_same_known_answers():
    # MATLAB: conv([1 2 3],[1 1],'same') = [3 5 3]; conv([1 2 3 4],[1 2 1],'same') = [4 8 12 11]
    assert np.array_equal(O.matlab_conv_same([1, 2, 3], [1, 1]), [3, 5, 3])
    assert np.array_equal(O.matlab_conv_same([1, 2, 3, 4], [1, 2, 1]), [4, 8, 12, 11])
    # the example of MATLAB's conv documentation (even-length kernel)
    assert np.array_equal(O.matlab_conv_same([-1, 2, 3, -2, 0, 1, 2], [2, 4, -1, 1]), [15, 5, -9, 7, 6, 7, -1])
    rng = np.random.default_rng(1)
    a = rng.standard_normal((5, 6, 7))
    k = rng.standard_normal(4)
    ref = np.apply_along_axis(O.matlab_conv_same, 1, a, k)
    assert np.allclose(O._conv_same_axis(a, k, 1), ref, atol=1e-12)


@pytest.mark.parametrize("F", [4, 5, 8, 9, 12])
def test_moving_average_matches_reflect_correlation(F):
    """BH_movingAverage_2 == separable correlation with the normalised
    Gaussian under symmetric (scipy 'reflect') boundaries; an even kernel is
    centred half a voxel high (MATLAB 'same'): scipy origin -1."""
    from scipy.ndimage import correlate1d
    rng = np.random.default_rng(F)
    a = rng.standard_normal((12, 14, 10))
    k = O.gaussian_taps(F)
    assert abs(k.sum() - 1) < 1e-12 and np.allclose(k, k[::-1])
    ref = a
    for ax in (2, 1, 0):
        ref = correlate1d(ref, k, axis=ax, mode="reflect", origin=-1 if F % 2 == 0 else 0)
    assert np.allclose(O.moving_average(a, F), ref, atol=1e-12)
    assert np.allclose(O.moving_average(np.full((8, 9, 10), 2.5), F), 2.5)


def test_local_normalize_invariance():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((10, 11, 12))
    assert np.allclose(O.local_normalize(3.0 * a - 7.0, 6), O.local_normalize(a, 6), atol=1e-9)


def test_preprocess_bandpass_half_grid_equals_full():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((12, 10, 16))
    bp = O.bandpass3d((16, 10, 12), 1e-3, 40.0, 8.0, 2.5)
    half = np.fft.irfftn(np.fft.rfftn(x) * bp[:, :, :9], s=x.shape, axes=(0, 1, 2))
    full = np.real(np.fft.ifftn(np.fft.fftn(x) * bp))
    assert np.allclose(half, full, atol=1e-12)


def test_mask3d_rectangle_edges():
    """size 40, radius 12: window 37, padVal low/high 2/1 plus the even-size
    shift -> 3/0; the window's taper times the volume-edge taper."""
    m = O.mask3d_rectangle((40, 36, 32), (12, 10, 8))
    t = 0.5 + 0.5 * np.cos(np.arange(7) * np.pi / 7)
    line = m[16, 18, :]
    assert np.all(line[:3] == 0)
    assert np.isclose(line[3], t[6] * t[3]) and np.isclose(line[6], t[3] * t[0])
    assert np.all(line[10:33] == 1.0)
    assert np.isclose(line[39], t[6] * t[6]) and np.isclose(line[33], t[0] * t[0])
    # separable: the volume is the outer product of its central lines
    lz, ly, lx = m[:, 18, 20], m[16, :, 20], m[16, 18, :]
    assert np.allclose(m, lz[:, None, None] * ly[None, :, None] * lx[None, None, :])


def test_tsearch_global_variance_squares_chunk_sums():
    rng = np.random.default_rng(6)
    tomo = rng.standard_normal((30, 36, 40))
    chunks, corners, gvar, mask = O.tsearch_preprocess(tomo, (24, 24, 24), (16, 16, 16), (2, 2, 1),
                                                       (4, 4, 2), (4, 4, 2), (20, 20, 20), (1e-3, 600.0, 28.0),
                                                       10.0, 8, 1)
    sums = []
    for (cx, cy, cz) in corners:
        t = np.pad(tomo, [(2, 2), (4, 4), (4, 4)], mode="symmetric")[cz:cz + 24, cy:cy + 24, cx:cx + 24]
        sums.append(O.tsearch_preprocess_chunk(t, (1e-3, 600.0, 28.0), 10.0, 8, -1.0, mask)[1])
    N = len(sums) * 24 ** 3
    assert np.isclose(gvar, sum(s * s for s in sums) / N - (sum(sums) / N) ** 2, rtol=1e-12)
    assert len(chunks) == 4 and corners == [(0, 0, 0), (0, 16, 0), (16, 0, 0), (16, 16, 0)]


# ------------------------------------------------ template-search peak picking

def test_mask3d_sphere_pins():
    """BH_mask3d('sphere'): the separable 5-tap convolution equals the full
    3-D convn with the BH_multi_gaussian3d kernel (scipy, zero boundary)
    applied to the shells; the core is 1; a voxel first inside at shell 1
    carries the recurrence w = t1, then w + (1 - w) t_s; symmetric about
    floor(N/2)."""
    from scipy.ndimage import convolve
    n, R = 31, 5.0
    m = O.mask3d_sphere((n, n, n), (R, R, R))
    c = n // 2
    assert m.shape == (n, n, n) and m[c, c, c] == 1.0 and m[c, c, c + 5] == 1.0
    z, y, x = np.meshgrid(*(np.arange(n) - c,) * 3, indexing="ij")
    r2 = x ** 2 + y ** 2 + z ** 2
    w = (r2 <= R * R).astype(float)
    t = 0.5 + 0.5 * np.cos(np.arange(1, 7) * np.pi / 7)
    for s in range(1, 7):
        w = w + ((r2 <= (R + s) ** 2).astype(float) - w) * t[s - 1]
    v = 0.0                      # r = 6: outside the core, inside from shell 1 on
    for s in range(1, 7):
        v = v + (1.0 - v) * t[s - 1]
    assert w[c, c, c + 6] == pytest.approx(v, abs=1e-15)
    g = O.gaussian3d_taps(5, 0.5)
    k3 = g[:, None, None] * g[None, :, None] * g[None, None, :]
    full = convolve(w, k3, mode="constant")
    full = full / full.max()
    full[r2 <= R * R] = 1.0
    full = O.pad_zeros3d(full, [0, 0, 0], [0, 0, 0], taper=True)
    assert np.allclose(m, full, atol=1e-12)
    assert np.allclose(m, m[::-1, ::-1, ::-1], atol=1e-12)
    assert np.all(m[:, :, 0] == 0) and m[c, c, c + 14] < 1e-6


def _planted_map(rng, shape=(40, 44, 48), centres=((20, 20, 22), (30, 12, 15), (12, 30, 33)),
                 amps=(5.0, 4.0, 3.0)):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    mag = 0.05 * rng.standard_normal(shape)
    for (cx, cy, cz), a in zip(centres, amps):
        mag += a * np.exp(-((x - cx - 0.3) ** 2 + (y - cy + 0.2) ** 2 + (z - cz) ** 2) / 2.0)
    ang = rng.integers(1, 20, size=shape).astype(np.int32)
    return mag, ang


def test_tsearch_peaks_recovers_planted_peaks():
    rng = np.random.default_rng(3)
    mag, ang = _planted_map(rng)
    rm = 4 + 7                           # rmDim = erase radius + 7, as the reference sizes it
    pad = lambda a: np.pad(a, rm)
    grid = np.stack([np.arange(19) * 10.0, np.full(19, 5.0), np.zeros(19)], 1)
    mask = O.mask3d_sphere((2 * rm + 1,) * 3, (4.0, 4.0, 4.0))
    pk = O.tsearch_peaks(pad(mag), pad(ang), grid, mask, rm, 0.98, 1.0, 10, 1000)
    assert len(pk) == 3                                  # the noise stays below the threshold
    assert np.all(np.diff(pk[:, 6]) < 0)                 # taken in order of score
    for p, (cx, cy, cz) in zip(pk, ((20, 20, 22), (30, 12, 15), (12, 30, 33))):
        # 1-based coordinates; the centre of mass pulls toward the +0.3/-0.2 offsets
        assert abs(p[0] - (cx + 1)) < 0.5 and abs(p[1] - (cy + 1)) < 0.5 and p[2] == pytest.approx(cz + 1, abs=0.05)
        zz, yy, xx = cz, cy, cx
        assert np.array_equal(p[3:6], grid[ang[zz, yy, xx] - 1])


def test_template_csv_row_format(tmp_path):
    from emclarity_amd.template_search import write_template_csv
    from emclarity_amd.geometry import BH_defineMatrix
    pk = np.array([[120.5, 64.0, 33.25, 57.2727272727, 12.0, -30.0, 7.891]])
    write_template_csv(tmp_path / "a.csv", tmp_path / "a.pos", pk, 2, n_previous=4)
    row = (tmp_path / "a.csv").read_text().split()
    assert len(row) == 26
    assert row[:10] == ["7.89", "2", "0", "5", "1", "1", "1", "1", "1", "0"]
    assert row[10:13] == ["120.500000", "64.000000", "33.250000"]
    assert row[13:16] == ["5.727273e+01", "12", "-30"]
    r = BH_defineMatrix(pk[0, 3:6], "Bah", "inv").reshape(-1, order="F")
    assert np.allclose([float(v) for v in row[16:25]], r, atol=1e-6) and row[25] == "1"
    assert (tmp_path / "a.pos").read_text() == "60.250000 32.000000 16.625000\n"


def _one_peak_map(head_id=0):
    """22^3 padded maps: a peak of 4 (id 5) at voxel (x, y, z) = (11, 10, 12)
    (0-based), 1.0 at x + 2 (id 5) and 0.5 at y - 1 (id 7) - both inside the
    5^3 box, only the latter inside the 3^3 one - every other voxel 0 with
    id 5; the first 343 ids of the map (its linear head, which the extra peaks
    read as Ang(cAng)) are head_id."""
    mag = np.zeros((22, 22, 22))
    ang = np.full((22, 22, 22), 5, np.int32)
    ang.reshape(-1)[:343] = head_id
    mag[12, 10, 11] = 4.0
    mag[12, 10, 13] = 1.0
    mag[12, 9, 11] = 0.5
    ang[12, 9, 11] = 7
    return mag, ang


def _one_peak_args():
    grid = np.stack([np.arange(9) * 10.0, np.full(9, 5.0), np.arange(9) * -2.0], 1)
    return grid, O.mask3d_sphere((7, 7, 7), (2.0, 2.0, 2.0))


def test_tsearch_peaks_multi_reduces_to_one_peak():
    rng = np.random.default_rng(3)
    mag, ang = _planted_map(rng)
    mag = mag.astype(np.float32)             # a single-precision map, as the reference's
    rm = 11
    pad = lambda a: np.pad(a, rm)
    grid = np.stack([np.arange(19) * 10.0, np.full(19, 5.0), np.zeros(19)], 1)
    mask = O.mask3d_sphere((2 * rm + 1,) * 3, (4.0, 4.0, 4.0))
    pk = O.tsearch_peaks(pad(mag), pad(ang), grid, mask, rm, 0.98, 1.0, 10, 1000)
    pm = O.tsearch_peaks_multi(pad(mag), pad(ang), grid, mask, rm, 0.98, 1.0, 10, 1000, 1, 2.0, None)
    assert pm.shape == (3, 10)
    assert np.allclose(pm[:, 0:3], 2.0 * pk[:, 0:3], rtol=0, atol=1e-12)
    assert np.array_equal(pm[:, 3:6], pk[:, 3:6]) and np.array_equal(pm[:, 9], pk[:, 6])
    assert not np.any(pm[:, 6:9])


def test_tsearch_peaks_multi_box_and_random_extras():
    """bDist = 1 + round(log(nPeaks)): 2 for 4 peaks, so the x + 2 voxel
    enters the centre of mass (x: 2 x 1 / 5.5, y: -0.5 / 5.5). The extra
    peaks: the id-7 voxel is the only non-zero value left, so peak 2 takes
    iSNR = 0.75 x 4 and (the map's head being id 0) a random angle; peaks 3
    and 4 find nothing left and keep that iSNR; nRandom runs 2, 3, 4."""
    mag, ang = _one_peak_map(0)
    grid, mask = _one_peak_args()
    normals = np.array([[[0.5, -1.0, 2.0], [1.5, 0.25, -0.5], [-2.0, 1.0, 0.125]]])
    pm = O.tsearch_peaks_multi(mag, ang, grid, mask, 3, 0.5, 1.0, 1, 100, 4, 2.0, normals)
    assert pm.shape == (1, 40)
    pos = 2.0 * np.array([12 + 2.0 / 5.5 - 3, 11 - 0.5 / 5.5 - 3, 13 - 3])
    assert np.allclose(pm[0, 0:3], pos, rtol=0, atol=1e-12)
    assert np.array_equal(pm[0, 3:6], grid[4]) and pm[0, 9] == 4.0
    for k, (nr, snr) in enumerate(zip((2, 3, 4), (3.0, 3.0, 3.0)), start=1):
        sub = pm[0, 10 * k:10 * k + 10]
        assert np.array_equal(sub[0:3], pm[0, 0:3])
        assert np.allclose(sub[3:6], pos + nr ** 2 * normals[0, k - 1], rtol=0, atol=1e-12)
        assert sub[9] == snr and not np.any(sub[6:9])
    # one peak: the 3^3 box leaves the x + 2 voxel out
    p1 = O.tsearch_peaks_multi(mag, ang, grid, mask, 3, 0.5, 1.0, 1, 100, 1, 1.0, None)
    assert np.allclose(p1[0, 0:3], [12 - 3, 11 - 0.5 / 4.5 - 3, 13 - 3], rtol=0, atol=1e-12)


def test_tsearch_peaks_multi_angles_from_the_id_map():
    """With a positive id at the box's linear index in the map's head (id 6),
    peak 2 takes ANGLE_LIST(6) and the mean of its class in the box (0.5, the
    one id-7 voxel); peak 3 then finds nothing left and is the first random
    one (nRandom = 2), keeping iSNR 0.5."""
    mag, ang = _one_peak_map(6)
    grid, mask = _one_peak_args()
    normals = np.array([[[0.5, -1.0, 2.0], [1.5, 0.25, -0.5]]])
    pm = O.tsearch_peaks_multi(mag, ang, grid, mask, 3, 0.5, 1.0, 1, 100, 3, 1.0, normals)
    assert np.array_equal(pm[0, 13:16], grid[5]) and pm[0, 19] == 0.5
    assert np.allclose(pm[0, 23:26], pm[0, 0:3] + 4 * normals[0, 1], rtol=0, atol=1e-12) and pm[0, 29] == 0.5


def test_template_csv_extra_peaks_on_one_line(tmp_path):
    from emclarity_amd.template_search import write_template_csv
    pk = np.array([[[120.5, 64.0, 33.25, 10.0, 12.0, -30.0, 7.891], [120.5, 64.0, 33.25, 1.5, 2.0, 3.0, 5.9]]])
    write_template_csv(tmp_path / "a.csv", tmp_path / "a.pos", pk, 2)
    lines = (tmp_path / "a.csv").read_text().splitlines()
    assert len(lines) == 1
    row = lines[0].split()
    assert len(row) == 52 and row[26] == "7.89" and row[39:42] == ["1.500000e+00", "2", "3"]


# ------------------------------------------------------- point-group operators

def _is_group(mats):
    for a in mats:
        assert abs(np.linalg.det(a) - 1.0) < 1e-12 and np.allclose(a @ a.T, np.eye(3), atol=1e-12)
        for b in mats[:12]:
            assert any(np.max(np.abs(a @ b - c)) < 1e-9 for c in mats)


@pytest.mark.parametrize("name,order", [("O", 24), ("I", 60), ("I2", 60)])
def test_point_groups(name, order):
    G = O.symmetry_matrices(name)
    assert len(G) == order
    _is_group(G)
    # 2-fold axes along x, y and z in every setting
    for a in np.eye(3):
        assert any(np.allclose(g, 2 * np.outer(a, a) - np.eye(3), atol=1e-12) for g in G)
    if name == "O":
        # 4-folds on the axes: the signed permutations with determinant +1
        assert all(np.allclose(np.sort(np.abs(g), axis=1)[:, -1], 1) for g in G)
    else:
        phi = (1 + 5 ** 0.5) / 2
        v = np.array([1.0, 0.0, phi]) if name == "I2" else np.array([0.0, 1.0, phi])
        v /= np.linalg.norm(v)
        # a 5-fold about the setting's vertex, and about none of the other setting's
        rot72 = [g for g in G if abs(np.trace(g) - (1 + 2 * np.cos(np.radians(72)))) < 1e-9]
        assert len(rot72) == 12
        assert any(np.allclose(g @ v, v, atol=1e-12) for g in rot72)
        w = np.array([0.0, 1.0, phi]) if name == "I2" else np.array([1.0, 0.0, phi])
        w /= np.linalg.norm(w)
        assert not any(np.allclose(g @ w, w, atol=1e-9) for g in rot72)


def test_dihedral_uses_the_cyclic_operators():
    """interpolator.m:199-210 keeps nSymMats = n for D<n>: the D4 average is
    the C4 average."""
    assert all(np.allclose(a, b) for a, b in zip(O.symmetry_matrices("D4"), O.symmetry_matrices("C4")))


def test_symmetrised_volume_is_invariant():
    rng = np.random.default_rng(5)
    v = rng.standard_normal((9, 9, 9))
    G = O.symmetry_matrices("O")
    a = O.interp3d(v, np.eye(3), [0, 0, 0], "Bah", "inv", "O")
    b = O.interp3d(v, G[7], [0, 0, 0], "Bah", "inv", "O")
    assert np.allclose(a, b, atol=1e-10)


# ------------------------------------------------------------- alignment loop

def test_align_grid_padded_agrees_with_unpadded():
    """The oracle's two-stage body with a window smaller than the calc size
    (the 63 -> 64 case of BH_multi_validArea, here 23 and 24 -> 26): the
    same particle seen through a smaller window gives the stage-1 shift of
    the W = C run to 0.05 px at the true orientation and at most others (a
    wrong orientation's peak is broad and may move), mean padding adds no
    edge, and stage 2 puts the truth first."""
    from emclarity_amd import synthetic as S
    truth = (40.0, 30.0, 20.0)
    pts = np.array([truth, (0.0, 0.0, 0.0), (60.0, 30.0, 20.0), (40.0, 45.0, 20.0), (40.0, 30.0, 40.0)])
    c = 26
    wedge = np.fft.fftshift(S.wedge_mask(c, 80.0))
    bp = O.bandpass3d((c, c, c), 1e-3, 40.0, 6.0, 2.5)
    pm = S.sphere_mask(c, c // 4).astype(np.float64)
    out = {}
    for w, pad in ((26, [[0] * 3, [0] * 3]), (23, [[1] * 3, [2] * 3]), (24, [[1] * 3, [1] * 3])):
        case = S.make_align_case(4, n=w, euler=truth, snr=50.0, n_blobs=10, blob_radius=5.0,
                                 shift=(1.0, -0.5, 0.25))
        vm = S.sphere_mask(w, 0.45 * w).astype(np.float64)
        ref = case.ref.astype(np.float64)
        inp = O.AlignInputs(particle=case.particle.astype(np.float64), ref=ref, ref_wgt_rot=None, ref_wgt=1.0,
                            ref_ft1=O.ref_ft1_of(ref, vm, pad), vol_mask=vm, peak_mask=pm, bandpass=bp,
                            wedge=wedge, angles=np.eye(3), shift_val=np.zeros(3), pad_calc=pad)
        out[w] = O.align_grid(inp, pts)
    for w in (23, 24):
        cc, est = out[w]
        assert int(np.argmax(cc)) == 0
        d = np.max(np.abs(est - out[26][1]), axis=1)
        assert d[0] < 0.05 and np.median(d) < 0.05, (w, d)


# ------------------------------------------------------------- median filter

@pytest.mark.parametrize("k", [3, 5, 7])
def test_median_filter3_brute_force(k, rng):
    """The oracle's medfilt3 (scipy 'reflect') against a direct statement:
    numpy's 'symmetric' pad (MATLAB's 'symmetric': the border sample
    repeated) and the median of every k^3 window; an isolated spike is
    removed and a constant volume is unchanged."""
    x = rng.standard_normal((7, 8, 9))
    h = k // 2
    p = np.pad(x, h, mode="symmetric")
    want = np.empty_like(x)
    for z in range(x.shape[0]):
        for y in range(x.shape[1]):
            for xx in range(x.shape[2]):
                want[z, y, xx] = np.median(p[z:z + k, y:y + k, xx:xx + k])
    assert np.array_equal(O.median_filter3(x, k), want)
    spike = np.zeros((7, 8, 9))
    spike[3, 4, 4] = 100.0
    assert not O.median_filter3(spike, k).any()
    assert np.array_equal(O.median_filter3(np.full((7, 8, 9), 2.5), k), np.full((7, 8, 9), 2.5))


# ------------------------------------------------------------ EMC_maskShape

def test_mask_shape_binary_sphere_and_shift():
    """The binary sphere (kernel false) against a brute-force FP64 count away
    from ties, and an integer shift is an exact translation."""
    m = O.mask_shape("sphere", (40, 40, 40), (9.3, 9.3, 9.3), (0, 0, 0), False)
    z, y, x = np.meshgrid(*(np.arange(40) - 20,) * 3, indexing="ij")
    assert np.array_equal(m, ((x * x + y * y + z * z) / 9.3 ** 2 <= 1).astype(float))
    s = O.mask_shape("sphere", (40, 40, 40), (9.3, 9.3, 9.3), (3, -2, 1), False)
    assert np.array_equal(s, np.roll(m, (1, -2, 3), axis=(0, 1, 2)))


def test_mask_shape_kernel_taps():
    """The default kernel: 9 taps for boxes under ~237 voxels, symmetric,
    sum 1, and the Gaussian is 0.5 at ceil(K/2)/2 from the centre."""
    k = O.mask_shape_kernel_taps((64, 64, 64), 0.04)
    assert len(k) == 9 and abs(float(k.sum()) - 1.0) < 1e-6 and np.allclose(k, k[::-1])
    assert len(O.mask_shape_kernel_taps((300, 300, 300), 0.04)) == 13   # round((12 - 1) / 2) 2 + 1
    middle = 2.5
    sigma = math.sqrt(-middle ** 2 / (2 * math.log(0.5)))
    assert abs(math.exp(-0.5 * (middle / sigma) ** 2) - 0.5) < 1e-12


def test_mask_shape_blurred_rectangle_is_separable():
    """A blurred rectangle is the outer product of its three blurred 1-D
    profiles (normalised, before the cut-off); the blurred sphere is 1 at its
    centre, exactly 0 beyond the enlarged radius plus the kernel's reach."""
    r = O.mask_shape("rectangle", (40, 36, 32), (8, 6, 5), (0, 0, 0), True)
    k = O.mask_shape_kernel_taps((40, 36, 32), 0.04).astype(np.float64)
    prof = []
    for n, rad in ((40, 8), (36, 6), (32, 5)):
        v = (np.abs(np.arange(n) - n // 2) <= rad + 4).astype(float)
        v[:5], v[-5:] = 0, 0
        v = np.convolve(v, k, mode="same")
        prof.append(v / v.max())
    outer = prof[2][:, None, None] * prof[1][None, :, None] * prof[0][None, None, :]
    keep = outer > 0.0011
    assert np.max(np.abs(r[keep] - outer[keep])) < 1e-12 and np.all(r[outer < 0.0009] == 0)
    s = O.mask_shape("sphere", (64, 64, 64), (20, 20, 20), (0, 0, 0), True)
    z, y, x = np.meshgrid(*(np.arange(64) - 32,) * 3, indexing="ij")
    assert s[32, 32, 32] == 1.0 and np.all(s[np.sqrt(x * x + y * y + z * z) > 25 + 4 * math.sqrt(3) + 1e-9] == 0)
