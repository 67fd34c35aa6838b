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
        self.bpVals = None

    def fwdFFT(self, inputVol, doNorm=0, doCenter=False, bp_vals=None):
        """fwdFFT(x) (:92-145): R2C half spectrum, unscaled. The normalise /
        centre / bandpass options of the reference are used only by the
        template-search preprocessing, which this package does not implement."""
        if doNorm or doCenter or bp_vals is not None:
            raise NotImplementedError("fwdFFT options (norm/center/bandpass) are not implemented")
        v = as_device(inputVol, np.float32, self.ctx)
        return self.plan.r2c(v)

    def invFFT(self, inputVol, doNorm=0):
        """invFFT(x) (:147-159), unscaled; the input is preserved."""
        if doNorm:
            raise NotImplementedError("invFFT normalisation is not implemented")
        v = as_device(inputVol, np.complex64, self.ctx).copy()
        return self.plan.c2r(v)
