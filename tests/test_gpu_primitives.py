"""GPU parity of every drop-in primitive against the FP64 oracle, through the
C ABI. Tolerances are stated per test; FP32 device arithmetic vs FP64."""
import math

import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.device import DeviceArray
from emclarity_amd.fourier import FFTPlan, fourierTransformer, mexFFT
from emclarity_amd.geometry import BH_defineMatrix
from emclarity_amd.interpolator import interpolator, mexXform3d
from emclarity_amd.masking import (BH_bandLimitCenterNormalize, BH_bandpass3d, BH_padZeros3d,
                                   BH_weightMaskMex, BH_weightMaskMex_batched, EMC_maskShape, mexCTF, mexSF3D)
from emclarity_amd.xcf import BH_multi_xcf_Rotational, BH_multi_xcf_Translational
from oracle import emc_oracle as O
from tests.conftest import load_tlt

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


# ------------------------------------------------------------ rotate (A1/A2)

@pytest.mark.parametrize("n,fwd", [(32, True), (32, False), (33, True), (20, False)])
def test_xform3d_matches_oracle(ctx, rng, n, fwd):
    v = rng.standard_normal((n, n + 2, n - 1)).astype(np.float32)
    M = BH_defineMatrix(rng.uniform(-180, 180, 3), "Bah", "forward")
    t = rng.uniform(-3, 3, 3)
    out = mexXform3d(v.shape[::-1], v, M, t, fwd).get()
    ref = O.xform3d(v, M.astype(np.float32).astype(np.float64), t.astype(np.float32), fwd)
    # FP32 coordinates: sample position error ~1e-5 px
    assert np.max(np.abs(out - ref)) < 2e-4


def test_xform3d_exact_cases(ctx, rng):
    v = rng.standard_normal((16, 16, 16)).astype(np.float32)
    # identity and integer shift are exact
    assert np.array_equal(mexXform3d((16, 16, 16), v, np.eye(3), [0, 0, 0], True).get(), v)
    s = mexXform3d((16, 16, 16), v, np.eye(3), [2, 0, 0], True).get()
    assert np.array_equal(s[:, :, 2:], v[:, :, :-2]) and not s[:, :, :2].any()
    # 90 degrees about z: exact permutation about n//2
    Rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    r = mexXform3d((16, 16, 16), v, Rz, [0, 0, 0], True).get()
    assert np.array_equal(r, O.xform3d(v, Rz, [0, 0, 0], True).astype(np.float32))


@pytest.mark.parametrize("symmetry", ["C3", "D4", "O", "I", "I2"])
def test_interpolator_symmetry(ctx, rng, symmetry):
    v = rng.standard_normal((24, 24, 24)).astype(np.float32)
    ip = interpolator(v, [0, 0, 0], [0, 0, 0], "Bah", "forward", "C1")
    ang, sh = [12.0, 33.0, -70.0], [0.5, -1.0, 0.25]
    out = ip.interp3d(ang, sh, "Bah", "inv", symmetry).get()
    ref = O.interp3d(v, ang, sh, "Bah", "inv", symmetry)
    assert np.max(np.abs(out - ref)) < 2e-4


# ------------------------------------------------------------ CTF / SF3D (A5-A7)

@pytest.mark.parametrize("half,centered,sq", [(False, True, True), (True, False, False),
                                              (False, False, True), (True, True, False)])
def test_mexctf_matches_oracle(ctx, half, centered, sq):
    args = dict(pixelSize=2.17, waveLength=0.01968697, CS=2.7, defocus1=31000.0, defocus2=27000.0,
                defocusAst=-60.0, AmpContrast=0.085)
    out = mexCTF(half, sq, 64, 48, centered=centered, **args).get()
    p = O.CtfParams(2.17, 0.01968697, 2.7, 0.085, 31000.0, 27000.0, -60.0)
    ref = O.ctf2d(64, 48, p, half=half, sq=sq, centered=centered, spacing_from_full_nx=True)
    assert out.shape == ref.shape
    # The phase is evaluated in FP32, as the reference's kernel does
    # (ctf.cu:45-64); it reaches ~200 rad at this grid's corner. Each of its
    # terms carries at most ~8 FP32 roundings (constants, k, r^2, r^4,
    # products, sums), so |d chi| <= 8 eps32 (|cs r^4| + |df r^2| (|d1| +
    # |d2|) + |amp|) per pixel; sin passes d chi through, its square twice
    # that; plus 2 ulp of sinf itself.
    cx = 64 // 2 + 1 if half else 64
    y, x = np.meshgrid(np.arange(48), np.arange(cx), indexing="ij")
    if centered:
        y = y - 24
        x = x if half else x - cx // 2
    else:
        y = np.where(y > 24, y - 48, y)
        x = x if half else np.where(x > cx // 2, x - cx, x)
    r2 = (x / (2.17 * 64)) ** 2 + (y / (2.17 * 48)) ** 2
    terms = p.cs_term * r2 ** 2 + p.df_term * r2 * (abs(p.dmean) + abs(p.ddiff)) + abs(p.amp_phase)
    bound = (2.0 if sq else 1.0) * 8 * np.finfo(np.float32).eps * terms + 2.4e-7
    assert np.all(np.abs(out - ref) <= bound), float(np.max(np.abs(out - ref) / bound))


def test_mexsf3d_matches_oracle(ctx):
    tilts = [dict(pixel=2.5, wavelength=0.019687, cs=2.7, df1=30000.0 + 50 * t, df2=29000.0 + 50 * t,
                  astig=15.0, amp=0.1, tilt=t, exposure=3.0 * abs(t) / 3 + 1.5,
                  occupancy=math.cos(math.radians(t))) for t in np.arange(-60, 61, 6.0)]
    n = 32
    keys = ["pixel", "wavelength", "cs", "df1", "df2", "astig", "amp", "tilt", "exposure", "occupancy"]
    arr = {k: np.array([t[k] for t in tilts], np.float32) for k in keys}
    out = mexSF3D(False, True, (n, n, n), arr["pixel"], arr["wavelength"], arr["cs"], arr["df1"],
                  arr["df2"], arr["astig"], arr["amp"], len(tilts), arr["tilt"], arr["exposure"],
                  arr["occupancy"]).get()
    ft = [{k: float(np.float32(t[k])) for k in keys} for t in tilts]
    ref = O.sf3d((n, n, n), ft, half=False, sq=True)
    assert rel(out, ref) < 1e-4


def test_weight_mask_on_reference_tlt(ctx):
    T = load_tlt("riboNew05_ali1_ctf.tlt")
    size, xyz, rg = (32, 32, 32), [812.4, 1101.0, 150.2], [[3838, 3710, 300], [0, 0, 0]]
    out = BH_weightMaskMex(size, 4, T, xyz, rg).get()
    ref = O.weight_mask_mex(size, 4, T, xyz, rg)
    assert rel(out, ref) < 1e-4
    assert abs(float(out.min()) - 1e-4) < 1e-6


def test_weight_mask_batched_matches_single_and_oracle(ctx):
    """Many particles, each with its own TLT (the reference's two fixtures and
    synthetic tables of other tilt ranges and defoci) and centre: the batched
    entry equals the single-particle entry bit for bit (same per-voxel
    arithmetic; the minimum is order-free), and a few against the oracle.
    110 particles of 128^3 with 31-47 tilts exceed one 256 MiB CTF stack, so
    the particles run in two chunks."""
    from emclarity_amd.synthetic import tilt_table
    rng = np.random.default_rng(12)
    fixtures = [load_tlt("riboNew05_ali1_ctf.tlt"), load_tlt("riboNew11_ali1_ctf.tlt")]
    tlts, xyzs = [], []
    for i in range(110):
        if i % 3 == 2:
            tlts.append(fixtures[i % 2])
        else:
            tlts.append(tilt_table(rng.uniform(45.0, 70.0), 3.0, rng.uniform(1.5, 4.0), 2.5))
        xyzs.append([rng.uniform(100, 3700), rng.uniform(100, 3600), rng.uniform(20, 280)])
    size, rg = (128, 128, 128), [[3838, 3710, 300], [0, 0, 0]]
    stack = BH_weightMaskMex_batched(size, 4, tlts, xyzs, rg)
    assert sum(len(t) for t in tlts) * 128 * 128 * 4 > (256 << 20)
    out = stack.get()
    for i in range(110):
        assert np.array_equal(out[i], BH_weightMaskMex(size, 4, tlts[i], xyzs[i], rg).get()), i
    for i in (0, 2, 57, 109):
        ref = O.weight_mask_mex(size, 4, tlts[i], xyzs[i], rg)
        assert rel(out[i], ref) < 1e-4, i
    small = BH_weightMaskMex_batched((33, 30, 28), 4, tlts[:3], xyzs[:3], rg).get()   # odd and uneven dims
    for i in range(3):
        assert np.array_equal(small[i], BH_weightMaskMex((33, 30, 28), 4, tlts[i], xyzs[i], rg).get())
    # The tables above leave the box's z-axis edge outside every slab, so
    # each volume's minimum is provably 0 and the batched entry adds the
    # 1e-4 floor at its store; a +-89 deg table reaches every voxel and
    # sends its chunk through the separate minimum pass. Both bit for bit.
    wide = [tilt_table(89.0, 3.0, 2.0, 2.5)] + tlts[:2]
    mixed = BH_weightMaskMex_batched((64, 64, 64), 4, wide, xyzs[:3], rg).get()
    for i in range(3):
        one = BH_weightMaskMex((64, 64, 64), 4, wide[i], xyzs[i], rg).get()
        assert np.array_equal(mixed[i], one), i
        assert abs(float(one.min()) - 1e-4) < 1e-6
    assert np.array_equal(BH_weightMaskMex_batched((64, 64, 64), 4, wide[1:], xyzs[1:3], rg).get(), mixed[1:])
    with pytest.raises(_lib.EmcError):
        _lib.call("emc_weight_mask_batched", ctx.handle, _lib.i3(size), 4.0, 1, None, None, None,
                  _lib.d_arr(rg[0]), _lib.d_arr(rg[1]), stack.vp)


@pytest.mark.parametrize("shape,size,radius,shift,kernel", [
    ("sphere", (64, 64, 64), (14, 14, 14), (0, 0, 0), False),          # peakMask, BH_alignRaw3d_v2.m:493
    ("sphere", (48, 40, 36), (12.5, 9.0, 7.5), (1.5, -2.0, 0.25), False),
    ("sphere", (64, 64, 64), (20, 20, 20), (0, 0, 0), True),           # volMask, :504
    ("sphere", (63, 50, 44), (14, 12, 9), (2.0, -1.0, 0.5), 0.1),
    ("cylinder", (48, 48, 40), (12, 10, 8), (1.0, 0.0, -2.0), True),
    ("rectangle", (40, 36, 32), (8, 6.4, 5), (2, -1, 0), True),
    ("rectangle", (40, 36, 32), (8, 6, 5), (0, 0, 0), False),
])
def test_emc_mask_shape_matches_oracle(ctx, shape, size, radius, shift, kernel):
    """EMC_maskShape against the oracle: binary shapes exactly (single-
    precision boundary test on both sides), blurred ones to 1e-5 except
    where the oracle sits within 1e-5 of the 0.001 cut-off."""
    out = EMC_maskShape(shape, size, radius, "gpu", {"shift": list(shift), "kernel": kernel}).get()
    ref = O.mask_shape(shape, size, radius, shift, kernel)
    assert out.shape == ref.shape
    if not kernel:
        assert np.array_equal(out, ref.astype(np.float32))
    else:
        near_cut = np.abs(ref - 0.001) < 1e-5
        assert np.max(np.abs(out - ref)[~near_cut]) < 1e-5
        assert float(out.max()) == 1.0


def test_emc_mask_shape_rejects_bad_arguments(ctx):
    with pytest.raises(_lib.EmcError):
        EMC_maskShape("sphere", (64, 64, 64), (1.0, 5, 5), "gpu", {})          # RADIUS must exceed 1
    with pytest.raises(_lib.EmcError):
        EMC_maskShape("sphere", (8, 64, 64), (5, 5, 5), "gpu", {"kernel": True})  # below the kernel's minimum SIZE
    with pytest.raises(ValueError):
        EMC_maskShape("sphere", (64, 64, 64), (5, 5, 5), "gpu", {"taper": 1})


# ------------------------------------------------------ spectral (A8-A12)

def test_bandpass_matches_oracle(ctx):
    out = BH_bandpass3d((48, 40, 32), 1e-3, 40.0, 8.0, "GPU", 2.5).get()
    ref = O.bandpass3d((48, 40, 32), 1e-3, 40.0, 8.0, 2.5)
    assert np.max(np.abs(out - ref)) < 1e-6
    half = BH_bandpass3d((48, 40, 32), 1e-3, 40.0, 8.0, "GPU", 2.5, half=True).get()
    assert np.array_equal(half, out[..., :25])


def test_pad_zeros3d_matches_oracle(ctx, rng):
    img = rng.standard_normal((20, 18, 22)).astype(np.float32)
    out = BH_padZeros3d(img, [2, -1, 3], [-2, 4, 0], "GPU", "single", 0.75).get()
    assert np.array_equal(out, O.pad_zeros3d(img, [2, -1, 3], [-2, 4, 0], 0.75).astype(np.float32))
    out = BH_padZeros3d(img, [1, 1, 1], [1, 1, 1], "GPU", "singleTaper", 0.0).get()
    assert np.max(np.abs(out - O.pad_zeros3d(img, [1, 1, 1], [1, 1, 1], 0.0, taper=True))) < 1e-6


def test_fft_roundtrip_and_numpy(ctx, rng):
    v = rng.standard_normal((24, 20, 32)).astype(np.float32)
    F, plan = mexFFT(DeviceArray.from_numpy(v), 2)
    Fh = F.get()
    ref = np.fft.rfftn(v.astype(np.float64))
    assert rel(Fh, ref) < 1e-5
    back, _ = mexFFT(F, 2, plan)
    assert rel(back.get() / v.size, v) < 1e-5
    ft = fourierTransformer(v)
    assert rel(ft.invFFT(ft.fwdFFT(v)).get() / v.size, v) < 1e-5
    full = DeviceArray.from_numpy(np.fft.fftn(v).astype(np.complex64))
    inv = plan.c2c(full, +1).get() / v.size
    assert np.max(np.abs(inv.real - v)) < 1e-4


@pytest.mark.parametrize("shape", [(12, 20, 32), (9, 14, 25)])
def test_fourier_transformer_options_match_oracle(ctx, rng, shape):
    """fwdFFT(x, doNorm, doCenter, bp, ctf) and invFFT(x, doNorm), even and
    odd nx; the bandpass is BH_bandpass3d's first nx/2+1 x-planes."""
    v = rng.standard_normal(shape).astype(np.float32) + 1.5
    nz, ny, nx = shape
    bpv = [1e-3, 40.0, 8.0, 2.5]
    bp_half = O.bandpass3d((nx, ny, nz), *bpv)[..., :nx // 2 + 1]
    ctf = rng.uniform(-1, 1, (nz, ny, nx // 2 + 1)).astype(np.float32)
    ft = fourierTransformer(v)
    for norm, centre, bp, c in ((1, False, None, None), (2, False, bpv, None), (0, True, bpv, None),
                                (2, True, None, ctf), (1, False, bpv, ctf)):
        out = ft.fwdFFT(v, norm, centre, bp, c).get()
        ref = O.fwd_fft(v, norm, centre, bp_half if bp else None, c)
        assert rel(out, ref) < 2e-5, (norm, centre, bp is not None, c is not None)
    F = ft.fwdFFT(v)
    for norm in (1, 2):
        back = ft.invFFT(F, norm).get()
        assert rel(back, v * v.size / np.sqrt(v.size) ** norm) < 1e-5


def test_band_limit_center_normalize_matches_oracle(ctx, rng):
    img = (rng.standard_normal((24, 24, 24)) + 2.0).astype(np.float32)
    bp = BH_bandpass3d((28, 28, 28), 1e-3, 40.0, 8.0, "GPU", 2.5)
    out = BH_bandLimitCenterNormalize(img, bp, "", [[2, 2, 2], [2, 2, 2]]).get()
    ref = O.band_limit_center_normalize(img, O.bandpass3d((28, 28, 28), 1e-3, 40.0, 8.0, 2.5),
                                        [[2, 2, 2], [2, 2, 2]])
    assert rel(out, ref[..., :15]) < 2e-5


# --------------------------------------------------------- correlation (A13/A14)

def test_xcf_translational_matches_oracle(ctx, rng):
    n = 32
    x = rng.standard_normal((n, n, n))
    ax = np.fft.fftfreq(n) * n
    z, y, xx = np.meshgrid(ax, ax, ax, indexing="ij")
    g = np.exp(-((xx - 2.3) ** 2 + (y + 1.6) ** 2 + (z - 0.4) ** 2) / (2 * 1.3 ** 2))
    A = np.fft.fftn(g + 0.01 * x).astype(np.complex64)
    B = np.ones((n, n, n), np.complex64)
    mask = O.xform3d(np.ones((n, n, n)), np.eye(3), [0, 0, 0], True)
    mask[:, :, : n // 2 - 6] = 0
    p, mx = BH_multi_xcf_Translational(A, B, mask.astype(np.float32), [3, 3, 3], return_max=True)
    ref, _, rmax = O.xcf_translational(A.astype(np.complex128), B.astype(np.complex128), mask,
                                       [3, 3, 3], return_map=True)
    assert np.max(np.abs(p - ref)) < 1e-4
    assert abs(mx - rmax) < 1e-5 * abs(rmax)
    # exact integer shift without COM
    v = rng.standard_normal((n, n, n))
    w = np.roll(v, (1, -4, 3), axis=(0, 1, 2))
    p = BH_multi_xcf_Translational(np.fft.fftn(w).astype(np.complex64),
                                   np.conj(np.fft.fftn(v)).astype(np.complex64), None, [0, 0, 0])
    assert np.array_equal(p, [3.0, -4.0, 1.0])


def test_xcf_translational_com_box_error(ctx):
    n = 16
    m = np.zeros((n, n, n))
    m[n // 2 - 1, 0, 0] = 1.0
    with pytest.raises(_lib.EmcError) as e:
        BH_multi_xcf_Translational(np.fft.fftn(m).astype(np.complex64), np.ones((n, n, n), np.complex64),
                                   None, [3, 3, 3])
    assert e.value.status == 5


def test_xcf_rotational_matches_oracle(ctx, rng):
    n = 24
    P = np.fft.fftn(rng.standard_normal((n, n, n))).astype(np.complex64)
    R = np.conj(np.fft.fftn(rng.standard_normal((n, n, n)) + 0.5 * np.fft.ifftn(P).real)).astype(
        np.complex64)
    wp = rng.uniform(0, 1, (n, n, n)).astype(np.float32)
    wr = rng.uniform(0, 1, (n, n, n)).astype(np.float32)
    c = BH_multi_xcf_Rotational(P, R, wp, wr)
    ref = O.xcf_rotational(P.astype(np.complex128), R.astype(np.complex128), wp, wr)
    assert abs(c - ref) < 1e-5 * max(abs(ref), 1e-3)
    assert abs(BH_multi_xcf_Rotational(P, np.conj(P), 1, 1) - 1.0) < 1e-6


def test_xcf_rotational_shells_matches_oracle(ctx, rng):
    """Shell-weighted CCC: radial shells (the layout calc_shells makes), NaN
    weight, abs flag, the ShellPlan reused across calls; an index beyond the
    volume is rejected."""
    from emclarity_amd.xcf import ShellPlan
    n = 24
    P = np.fft.fftn(rng.standard_normal((n, n, n))).astype(np.complex64)
    R = np.conj(np.fft.fftn(rng.standard_normal((n, n, n)) + 0.5 * np.fft.ifftn(P).real)).astype(np.complex64)
    wp = rng.uniform(0, 1, (n, n, n)).astype(np.float32)
    wr = rng.uniform(0, 1, (n, n, n)).astype(np.float32)
    f = np.fft.fftfreq(n) * n
    z, y, x = np.meshgrid(f, f, f, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z).ravel()
    edges = [0, 2.5] + list(np.arange(3.5, 12.5, 1.5))
    shells = [np.flatnonzero((lo <= r) & (r < hi)).astype(np.uint32) + 1 for lo, hi in zip(edges[:-1], edges[1:])]
    w = float(sum(len(s) for s in shells[1:]))
    P64, R64 = P.astype(np.complex128), R.astype(np.complex128)
    for flag in (1, 0):
        cell = [flag, w] + shells
        plan = ShellPlan(cell)
        for Wp, Wr in ((wp, wr), (wp, np.full((n, n, n), np.nan, np.float32)), (1, 1)):
            ref = O.xcf_rotational(P64, R64, Wp, Wr, cell)
            for c in (BH_multi_xcf_Rotational(P, R, Wp, Wr, plan), BH_multi_xcf_Rotational(P, R, Wp, Wr, cell)):
                assert abs(c - ref) < 1e-5 * max(abs(ref), 1e-6), (flag, c, ref)
        plan.close()
    with pytest.raises(_lib.EmcError):
        BH_multi_xcf_Rotational(P, R, 1, 1, [1, 1.0, np.array([n ** 3 + 1])])


def test_new_entry_points_reject_bad_arguments(ctx):
    """Error paths: fwdFFT option values, 0-based shell indices, a padCalc
    that does not fit the calc box; each raises EmcError with the library
    state intact (a good call follows)."""
    import ctypes as C
    from emclarity_amd.xcf import ShellPlan
    F = DeviceArray((8, 8, 5), np.complex64)
    for bad in ((3, 0), (-1, 0)):
        with pytest.raises(_lib.EmcError):
            _lib.call("emc_ft_condition", ctx.handle, F.vp, _lib.i3((8, 8, 8)), bad[0], bad[1], None, None)
    with pytest.raises(ValueError):           # the host mirror checks first
        ShellPlan([1, 1.0, np.array([0, 1, 2])])
    off = np.array([0, 3], np.int64)
    idx = np.array([0, 1, 2], np.uint32)
    h = C.c_void_p()
    with pytest.raises(_lib.EmcError):        # and the C ABI on its own
        _lib.call("emc_shell_plan_create", ctx.handle, 1, 1.0, 1, off.ctypes.data_as(C.POINTER(C.c_int64)),
                  idx.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(h))
    v = DeviceArray.from_numpy(np.ones((8, 8, 8), np.float32))
    big = DeviceArray.from_numpy(np.ones((10, 10, 10), np.float32))
    z = (C.c_void_p * 1)(big.ptr)
    ccc = np.zeros(1)
    est = np.zeros(3)
    dp = C.POINTER(C.c_double)
    with pytest.raises(_lib.EmcError):   # window 8 with 3 planes below leaves -1 above in 10
        _lib.call("emc_align_grid_padded", ctx.handle, 8, 10, _lib.i3((3, 3, 3)), 1, None, 1, v.vp, v.vp, None,
                  None, z, None, None, big.vp, z, big.vp, _lib.d_arr(np.eye(3).ravel()), _lib.d_arr([0, 0, 0]),
                  _lib.i3((1, 1, 1)), np.zeros(3).ctypes.data_as(dp), 1, 3, ccc.ctypes.data_as(dp),
                  est.ctypes.data_as(dp))
    zv = (C.c_void_p * 1)(v.ptr)
    for mode in (1, 3):                  # the translation stage without a peak mask
        with pytest.raises(_lib.EmcError, match="peak mask"):
            _lib.call("emc_align_grid_padded", ctx.handle, 8, 8, _lib.i3((0, 0, 0)), 1, None, 1, v.vp, v.vp, None,
                      None, zv, None, None, None, zv, v.vp, _lib.d_arr(np.eye(3).ravel()), _lib.d_arr([0, 0, 0]),
                      _lib.i3((1, 1, 1)), np.zeros(3).ctypes.data_as(dp), 1, mode, ccc.ctypes.data_as(dp),
                      est.ctypes.data_as(dp))
    _lib.call("emc_ft_condition", ctx.handle, F.vp, _lib.i3((8, 8, 8)), 1, 0, None, None)


def test_device_block_cache(ctx):
    """emc_free keeps emc_malloc blocks for reuse (same size class, same
    address), data written after reuse is intact, and a second free of the
    same block is rejected."""
    import ctypes as C
    a = DeviceArray.from_numpy(np.arange(1000, dtype=np.float32))
    p = a.ptr
    a.free()
    b = DeviceArray.from_numpy(np.full(900, 7.0, np.float32))   # same 512-byte class
    assert b.ptr == p
    assert np.array_equal(b.get(), np.full(900, 7.0, np.float32))
    big = DeviceArray.from_numpy(np.ones(3 << 20, np.float32))
    q = big.ptr
    big.free()
    c = DeviceArray(((3 << 20) - 100,), np.float32)                # same 2 MiB class
    assert c.ptr == q
    c.free()
    with pytest.raises(_lib.EmcError):
        _lib.call("emc_free", ctx.handle, C.c_void_p(q))
