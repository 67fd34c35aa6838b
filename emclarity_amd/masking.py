"""Weights and spectral helpers: BH_bandpass3d, BH_padZeros3d,
BH_bandLimitCenterNormalize, mexSF3D, BH_weightMaskMex, mexCTF."""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceArray, as_device, default_context
from .fourier import FFTPlan


def BH_bandpass3d(SIZE, HIGH_THRESH, HIGH_CUT, LOW_CUT, METHOD="GPU", PIXEL_SIZE=1.0, half=False,
                  ctx=None) -> DeviceArray:
    """masking/BH_bandpass3d.m (numeric pixel size): full grid, FFT order,
    shape (nz, ny, nx) (or the first nx//2+1 x-planes with half=True)."""
    ctx = ctx or default_context()
    nx, ny, nz = (list(SIZE) + [1])[:3]
    cx = nx // 2 + 1 if half else nx
    out = DeviceArray((nz, ny, cx), np.float32, ctx)
    _lib.call("emc_bandpass3d", ctx.handle, _lib.i3((nx, ny, nz)), int(bool(half)),
              float(HIGH_THRESH), float(HIGH_CUT), float(LOW_CUT), float(PIXEL_SIZE), out.vp)
    return out


def BH_padZeros3d(IMAGE, PADLOW, PADTOP, METHOD="GPU", PRECISION="single", extrapVal=0.0,
                  ctx=None) -> DeviceArray:
    """masking/BH_padZeros3d.m: pads (x,y,z) may be negative (trim);
    PRECISION 'singleTaper' applies the 7-voxel cosine edge."""
    if PRECISION not in ("single", "singleTaper"):
        raise ValueError(f"PRECISION must be single or singleTaper, not {PRECISION}")
    v = as_device(IMAGE, np.float32, ctx)
    lo = [int(x) for x in PADLOW]
    hi = [int(x) for x in PADTOP]
    w = [v.dims[k] - max(-lo[k], 0) - max(-hi[k], 0) for k in range(3)]
    o = [w[k] + max(lo[k], 0) + max(hi[k], 0) for k in range(3)]
    out = DeviceArray((o[2], o[1], o[0]), np.float32, v.ctx)
    _lib.call("emc_pad_zeros3d", v.ctx.handle, v.vp, _lib.i3(v.dims), _lib.i3(lo), _lib.i3(hi),
              float(extrapVal), int(PRECISION == "singleTaper"), out.vp)
    return out


def BH_bandLimitCenterNormalize(IMG, BANDPASS, MASK, PAD, PRECISION="single",
                                plan: FFTPlan | None = None, ctx=None) -> DeviceArray:
    """masking/BH_bandLimitCenterNormalize.m on the half spectrum.
    BANDPASS: DeviceArray / array (full grid FFT order of the padded size),
    or None / '' for none. PAD: [[lowx,lowy,lowz],[topx,topy,topz]]."""
    v = as_device(IMG, np.float32, ctx)
    pad = np.asarray(PAD, dtype=int).reshape(2, 3)
    pdims = tuple(int(v.dims[k] + pad[0, k] + pad[1, k]) for k in range(3))
    plan = plan or FFTPlan(pdims, 1, v.ctx)
    bp = None
    if BANDPASS is not None and not isinstance(BANDPASS, str):
        if np.isscalar(BANDPASS):
            bp = DeviceArray.from_numpy(np.full(pdims[::-1], BANDPASS, np.float32), ctx=v.ctx)
        else:
            bp = as_device(BANDPASS, np.float32, v.ctx)
    out = DeviceArray((pdims[2], pdims[1], pdims[0] // 2 + 1), np.complex64, v.ctx)
    _lib.call("emc_band_limit_center_normalize", v.ctx.handle, plan.handle, v.vp, _lib.i3(v.dims),
              _lib.i3(pad[0]), _lib.i3(pad[1]), bp.vp if bp is not None else None, out.vp)
    return out


def mexSF3D(doHalfGrid, doSqCTF, wantedSize, pixelSize, waveLength, CS, defocus1, defocus2,
            defocusAst, AmpContrast, nTilts, tiltAngle, exposure, occupancy, launch=1,
            ctx=None) -> DeviceArray:
    """SF3D = mexSF3D(...) (mexFiles/mexSF3D.cu:85-303); per-tilt arrays."""
    ctx = ctx or default_context()
    nx, ny, nz = (int(s) for s in wantedSize)
    n = int(nTilts)
    arrs = [np.asarray(a, np.float32).ravel() for a in (pixelSize, waveLength, CS, defocus1,
                                                        defocus2, defocusAst, AmpContrast,
                                                        tiltAngle, exposure, occupancy)]
    for a in arrs:
        if a.size < n:
            raise ValueError("per-tilt arrays shorter than nTilts")
    out = DeviceArray((nz, ny, nx), np.float32, ctx)
    _lib.call("emc_sf3d", ctx.handle, int(bool(doHalfGrid)), int(bool(doSqCTF)),
              _lib.i3((nx, ny, nz)), n, *[_lib.f_arr(a[:n]) for a in arrs], out.vp)
    return out


def BH_weightMaskMex(SIZE, SAMPLING, TLT, xyzSubTomo, reconGeometry, ctx=None) -> DeviceArray:
    """masking/BH_weightMaskMex.m: TLT (nTilts x 22) -> SF3D - (min - 1e-4).
    reconGeometry = [[size x,y,z],[shift x,y,z]]."""
    ctx = ctx or default_context()
    nx, ny, nz = (int(s) for s in SIZE)
    T = np.ascontiguousarray(np.asarray(TLT, dtype=np.float64))
    if T.ndim != 2 or T.shape[1] < 22:
        raise ValueError("TLT must have 22 columns")
    T = np.ascontiguousarray(T[:, :22])
    rg = np.asarray(reconGeometry, dtype=np.float64).reshape(2, 3)
    out = DeviceArray((nz, ny, nx), np.float32, ctx)
    _lib.call("emc_weight_mask", ctx.handle, _lib.i3((nx, ny, nz)), float(SAMPLING),
              _lib.d_arr(T.ravel()), T.shape[0], _lib.d_arr(np.asarray(xyzSubTomo, float).ravel()),
              _lib.d_arr(rg[0]), _lib.d_arr(rg[1]), out.vp)
    return out


def BH_weightMaskMex_batched(SIZE, SAMPLING, TLTs, xyzSubTomos, reconGeometry, out=None,
                             ctx=None) -> DeviceArray:
    """BH_weightMaskMex for many particles of one box size in one call
    (emc_weight_mask_batched): TLTs, a list of per-particle TLT tables
    (nTilts_p x 22), xyzSubTomos (n, 3) the particle centres, one shared
    reconGeometry. Returns the (n, nz, ny, nx) stack of SF3D - (min - 1e-4)
    (into ``out`` when given)."""
    ctx = ctx or (out.ctx if out is not None else default_context())
    nx, ny, nz = (int(s) for s in SIZE)
    tabs = [np.asarray(T, dtype=np.float64) for T in TLTs]
    for T in tabs:
        if T.ndim != 2 or T.shape[1] < 22 or T.shape[0] < 1:
            raise ValueError("every TLT must have 22 columns and at least one row")
    n = len(tabs)
    xyz = np.ascontiguousarray(np.asarray(xyzSubTomos, dtype=np.float64).reshape(n, 3))
    rows = np.ascontiguousarray(np.concatenate([T[:, :22] for T in tabs], axis=0)) if n else np.zeros((0, 22))
    counts = np.ascontiguousarray([T.shape[0] for T in tabs], dtype=np.int32)
    rg = np.asarray(reconGeometry, dtype=np.float64).reshape(2, 3)
    shape = (n, nz, ny, nx)
    if out is None:
        out = DeviceArray(shape, np.float32, ctx)
    elif out.shape != shape or out.dtype != np.float32:
        raise ValueError(f"out must be float32 {shape}")
    _lib.call("emc_weight_mask_batched", ctx.handle, _lib.i3((nx, ny, nz)), float(SAMPLING), n,
              rows.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)),
              counts.ctypes.data_as(_lib.C.POINTER(_lib.C.c_int32)),
              xyz.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)), _lib.d_arr(rg[0]), _lib.d_arr(rg[1]), out.vp)
    return out


def mexCTF(doHalfGrid, doSqCTF, nX, nY, pixelSize, waveLength, CS, defocus1, defocus2, defocusAst,
           AmpContrast, centered=False, ctx=None) -> DeviceArray:
    """ctf = mexCTF(...) (mexFiles/mexCTF.cu:1-121): shape (nY, nx_out)."""
    ctx = ctx or default_context()
    nx, ny = int(nX), int(nY)
    cx = nx // 2 + 1 if doHalfGrid else nx
    out = DeviceArray((ny, cx), np.float32, ctx)
    _lib.call("emc_ctf2d", ctx.handle, int(bool(doHalfGrid)), int(bool(doSqCTF)), nx, ny,
              float(pixelSize), float(waveLength), float(CS), float(defocus1), float(defocus2),
              float(defocusAst), float(AmpContrast), int(bool(centered)), out.vp)
    return out


def BH_mask3d(SHAPE, SIZE, RADIUS, CENTER, ctx=None) -> DeviceArray:
    """masking/BH_mask3d.m at CENTER [0,0,0] for the shapes template search
    uses. 'rectangle' (the valid-area mask, BH_templateSearch3d_2.m:402): a
    window of ones 2*RADIUS+13 per axis (:102-108), padded to SIZE with the
    7-voxel cosine taper and the half-voxel origin shift of even sizes
    (:244-280), then tapered at the volume edge (:462-466). 'sphere' (the
    peak erase mask, :1102): emc_mask3d_sphere. Other shapes are not part of
    this path."""
    shape = str(SHAPE).lower()
    if shape not in ("rectangle", "sphere") or any(float(c) != 0.0 for c in CENTER):
        raise NotImplementedError("BH_mask3d: only 'rectangle' and 'sphere' at [0,0,0] are provided")
    if shape == "sphere":
        from .device import default_context
        c = ctx or default_context()
        size = [int(v) for v in SIZE]
        out = DeviceArray((size[2], size[1], size[0]), np.float32, c)
        _lib.call("emc_mask3d_sphere", c.handle, _lib.i3(size), _lib.d_arr(np.asarray(RADIUS, np.float64)), out.vp)
        return out
    from .sizing import BH_multi_padVal
    size = np.asarray(SIZE, dtype=np.int64)
    win = (2 * np.asarray(RADIUS, dtype=np.float64) + 13).astype(np.int64)
    pv = BH_multi_padVal(win, size)
    shift = 1 - size % 2
    pre = pv[0] + shift
    post = pv[1] - shift
    ones = DeviceArray.from_numpy(np.ones(tuple(int(w) for w in win[::-1]), np.float32), ctx=ctx)
    # negative pads trim (BH_padZeros3d), as the reference trims before padding
    w = BH_padZeros3d(ones, np.minimum(pre, 0), np.minimum(post, 0), ctx=ctx)
    w = BH_padZeros3d(w, np.maximum(pre, 0), np.maximum(post, 0), "GPU", "singleTaper", ctx=ctx)
    return BH_padZeros3d(w, [0, 0, 0], [0, 0, 0], "GPU", "singleTaper", ctx=ctx)


def BH_movingAverage_2(img, featureSize, ctx=None) -> DeviceArray:
    """statistics/BH_movingAverage_2.m: Gaussian (sigma 3, featureSize taps)
    moving average with symmetric padding (emc_moving_average)."""
    v = as_device(img, np.float32, ctx)
    out = DeviceArray(v.shape, np.float32, v.ctx)
    _lib.call("emc_moving_average", v.ctx.handle, v.vp, _lib.i3(v.dims), int(featureSize), out.vp)
    return out


_MASK_SHAPES = {"sphere": 0, "cylinder": 1, "rectangle": 2}


def EMC_maskShape(SHAPE, SIZE, RADIUS, METHOD="gpu", OPTION=None, ctx=None) -> DeviceArray:
    """MASK = EMC_maskShape(SHAPE, SIZE, RADIUS, METHOD, OPTION)
    (masking/EMC_maskShape.m) for 3-D SIZE [x, y, z]: SHAPE 'sphere',
    'cylinder' or 'rectangle'; OPTION a dict or a cell-like list of (field,
    value) pairs with 'shift' ([x, y, z], default 0) and 'kernel' (True = the
    default 0.04, False / 0 = binary, or a fraction in (0, 1]); origin 1 and
    single precision, the reference's defaults. METHOD is accepted for the
    signature: the mask is built on the device. Returns (nz, ny, nx)."""
    from .device import default_context
    opt = dict(OPTION or {}) if not isinstance(OPTION, (list, tuple)) else {k: v for k, v in OPTION}
    unknown = set(opt) - {"shift", "kernel", "origin", "precision"}
    if unknown:
        raise ValueError(f"EMC_maskShape: unknown option(s) {sorted(unknown)}")
    if opt.get("origin", 1) != 1 or str(opt.get("precision", "single")).lower() != "single":
        raise NotImplementedError("EMC_maskShape: origin 1 and single precision (the reference defaults)")
    shape = str(SHAPE).lower()
    if shape not in _MASK_SHAPES:
        raise ValueError("SHAPE should be 'sphere', 'cylinder' or 'rectangle'")
    size = [int(v) for v in SIZE]
    if len(size) != 3:
        raise NotImplementedError("EMC_maskShape: 3-D SIZE only")
    kernel = opt.get("kernel", 0.04)
    kernel = 0.04 if kernel is True else (0.0 if kernel is False else float(kernel))
    shift = np.asarray(opt.get("shift", (0.0, 0.0, 0.0)), dtype=np.float64).reshape(3)
    c = ctx or default_context()
    out = DeviceArray((size[2], size[1], size[0]), np.float32, c)
    _lib.call("emc_mask_shape", c.handle, _MASK_SHAPES[shape], _lib.i3(size),
              _lib.d_arr(np.asarray(RADIUS, np.float64).reshape(3)), _lib.d_arr(shift), kernel, out.vp)
    return out
