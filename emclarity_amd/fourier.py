"""mexFFT plans and the fourierTransformer class (testScripts/fourierTransformer.m)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray, as_device, default_context


class FFTPlan:
    """mexFFT plan pair (mexFiles/mexFFT.cu:178-210) for real dims (nx,ny,nz)."""

    def __init__(self, dims, batch=1, ctx=None):
        self.ctx = ctx or default_context()
        self.dims = tuple(int(d) for d in dims)
        self.batch = int(batch)
        h = C.c_void_p()
        _lib.call("emc_fft_plan_create", self.ctx.handle, _lib.i3(self.dims), self.batch, C.byref(h))
        self.handle = h

    @property
    def half_shape(self):
        nx, ny, nz = self.dims
        return (nz, ny, nx // 2 + 1)

    def r2c(self, x: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
        nx, ny, nz = self.dims
        out = out or DeviceArray((self.batch,) + self.half_shape if self.batch > 1 else
                                 self.half_shape, np.complex64, self.ctx)
        _lib.call("emc_fft_r2c", self.handle, x.vp, out.vp)
        return out

    def c2r(self, x: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
        """Half spectrum -> real; the input is overwritten (as cuFFT C2R)."""
        nx, ny, nz = self.dims
        out = out or DeviceArray((self.batch, nz, ny, nx) if self.batch > 1 else (nz, ny, nx),
                                 np.float32, self.ctx)
        _lib.call("emc_fft_c2r", self.handle, x.vp, out.vp)
        return out

    def c2c(self, x: DeviceArray, direction: int, out: DeviceArray | None = None) -> DeviceArray:
        nx, ny, nz = self.dims
        out = out or DeviceArray((nz, ny, nx), np.complex64, self.ctx)
        _lib.call("emc_fft_c2c", self.handle, x.vp, out.vp, int(direction))
        return out

    def destroy(self):
        if self.handle:
            _lib.call("emc_fft_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def mexFFT(x, invTrim, plan: FFTPlan | None = None, ctx=None):
    """ft = mexFFT(x, invTrim[, plan]): real input -> R2C half spectrum;
    complex (half) input -> C2R of width 2*nx_half - invTrim (mexFFT.cu:98-128).
    Returns (result, plan). No scaling."""
    xd = as_device(x, None, ctx) if isinstance(x, DeviceArray) else \
        DeviceArray.from_numpy(np.asarray(x), ctx=ctx)
    if xd.dtype == np.complex64:
        nh, ny, nz = xd.dims
        dims = (2 * nh - int(invTrim), ny, nz)
        plan = plan or FFTPlan(dims, 1, xd.ctx)
        tmp = xd.copy()            # C2R overwrites its input
        return plan.c2r(tmp), plan
    xd = xd if xd.dtype == np.float32 else DeviceArray.from_numpy(xd.get().astype(np.float32))
    plan = plan or FFTPlan(xd.dims, 1, xd.ctx)
    return plan.r2c(xd), plan


class fourierTransformer:
    """R2C/C2R with persistent plans (fourierTransformer.m:38-159)."""

    def __init__(self, inputVol, ctx=None):
        v = as_device(inputVol, np.float32, ctx)
        self.ctx = v.ctx
        self.inputSize = v.dims
        self.halfDimSize = self.inputSize[0] // 2 + 1
        self.invTrim = 1 if self.inputSize[0] % 2 else 2          # :62-66
        self.normalization_factor = 1.0 / np.sqrt(np.prod(self.inputSize))
        self.plan = FFTPlan(self.inputSize, 1, self.ctx)
        self.bandpass = None
        self.bpVals = [0.0, 0.0, 0.0, 0.0]

    def makeBandPass(self, sizeInput, bpValsNew):
        """(:294-328) BH_bandpass3d(size, bp(1), bp(2), bp(3), 'GPU', bp(4))
        cut to the first nx/2+1 x-planes; rebuilt only when
        |sum(bpVals - bpValsNew)| > 1e-4, as the reference compares."""
        bp = [float(v) for v in bpValsNew]
        if self.bandpass is None or abs(sum(np.subtract(self.bpVals, bp))) > 1e-4:
            from .masking import BH_bandpass3d
            self.bandpass = BH_bandpass3d(sizeInput, bp[0], bp[1], bp[2], "GPU", bp[3], half=True, ctx=self.ctx)
            self.bpVals = bp

    def fwdFFT(self, inputVol, doNorm=0, doCenter=False, bp_vals=None, ctf=None):
        """fwdFFT(x[, doNorm, doCenter[, bp[, ctf]]]) (:92-145): R2C half
        spectrum; then x ctf (half grid), x (1/sqrt(numel))^doNorm unless
        centring, x bandpass (bp_vals = [HIGH_THRESH, HIGH_CUT, LOW_CUT,
        PIXEL_SIZE]), and with doCenter F(0) = 0 and F / sqrt(2 sum|F|^2)
        over the x-planes below nx/2+1-invTrim."""
        v = as_device(inputVol, np.float32, self.ctx)
        ft = self.plan.r2c(v)
        bp = None
        if bp_vals is not None and not (np.isscalar(bp_vals) or len(bp_vals) == 0):
            self.makeBandPass(v.dims, bp_vals)
            bp = self.bandpass
        c = None if ctf is None else as_device(ctf, np.float32, self.ctx)
        if c is not None and c.shape != ft.shape:
            raise ValueError(f"ctf must be the half grid {ft.shape}, not {c.shape}")
        if doNorm or doCenter or bp is not None or c is not None:
            _lib.call("emc_ft_condition", self.ctx.handle, ft.vp, _lib.i3(self.inputSize), int(doNorm),
                      int(bool(doCenter)), bp.vp if bp is not None else None, c.vp if c is not None else None)
        return ft

    def invFFT(self, inputVol, doNorm=0):
        """invFFT(x[, doNorm]) (:147-159): C2R, then x (1/sqrt(numel))^doNorm;
        the input is preserved."""
        v = as_device(inputVol, np.complex64, self.ctx).copy()
        out = self.plan.c2r(v)
        if doNorm:
            _lib.call("emc_scale", self.ctx.handle, out.vp, C.c_int64(int(np.prod(out.shape))),
                      float(self.normalization_factor ** int(doNorm)))
        return out
