"""mexXform3d and the interpolator class (testScripts/interpolator.m).

``interpolator`` keeps the input volume resident on the device (the
reference rebuilds a CUDA texture per call, interpolator.m:123-151) and
resamples it with emc_interp3d: trilinear, zero outside, symmetry mates
summed and divided by nSym.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceArray, as_device
from .geometry import BH_defineMatrix, symmetry_matrices


def mexXform3d(size, vol, R9, t, doFwd, ctx=None) -> DeviceArray:
    """[out] = mexXform3d(size, vol, single R[9], single t[3], doFwd)
    (mexFiles/mexXform3d.cu:93-324). R9 is the MATLAB column-major 9-vector."""
    v = as_device(vol, np.float32, ctx)
    dims = tuple(int(s) for s in size)
    if tuple(v.dims) != dims:
        raise ValueError(f"size {dims} does not match the volume {v.dims}")
    out = DeviceArray(v.shape, np.float32, v.ctx)
    R = np.asarray(R9, dtype=np.float32).reshape(-1, order="F") if np.ndim(R9) == 2 else \
        np.asarray(R9, dtype=np.float32).ravel()
    _lib.call("emc_xform3d", v.ctx.handle, v.vp, _lib.i3(dims), _lib.f_arr(R),
              _lib.f_arr(np.asarray(t, np.float32).ravel()), int(bool(doFwd)), out.vp)
    return out


class interpolator:
    """interpolator(inputVol, angles, shifts, convention, direction, symmetry)
    (interpolator.m:29-83); ``interp3d`` (:99-154)."""

    def __init__(self, inputVol, angles=(0, 0, 0), shifts=(0, 0, 0), convention="Bah",
                 direction="forward", symmetry="C1", ctx=None):
        self.input_volume = as_device(inputVol, np.float32, ctx)
        self.input_size = self.input_volume.dims
        self.symmetry_type = ""
        self.resampledVol = self.interp3d(angles, shifts, convention, direction, symmetry)

    def _check_symmetry(self, symmetry, convention):
        if symmetry != self.symmetry_type:
            if symmetry[0] != "C" and convention != "Bah":
                raise ValueError("Alternate conventions like Helical only support CX symmetry")
            self.symmetry_matrices = symmetry_matrices(symmetry, convention)
            self.nSymMats = len(self.symmetry_matrices)
            self.symmetry_type = symmetry

    def interp3d(self, angles, shifts, convention="Bah", direction="forward", symmetry="C1"):
        self._check_symmetry(symmetry, convention)
        a = np.asarray(angles, dtype=np.float64)
        if a.size == 3:
            M = BH_defineMatrix(a.ravel(), convention, direction)    # :159-161
        else:
            M = a.reshape(3, 3, order="F") if a.ndim == 1 else a.reshape(3, 3)
        if len(np.asarray(shifts).ravel()) != 3:
            raise ValueError("Shifts must have 3 elements")
        d = direction.lower()
        if d not in ("forward", "inv"):
            raise ValueError(f"Did not understand the transform direction {direction}")
        S = np.stack([s.reshape(-1, order="F") for s in self.symmetry_matrices]).ravel()
        out = DeviceArray(self.input_volume.shape, np.float32, self.input_volume.ctx)
        _lib.call("emc_interp3d", out.ctx.handle, self.input_volume.vp, _lib.i3(self.input_size),
                  _lib.d_arr(M.reshape(-1, order="F")), len(self.symmetry_matrices), _lib.d_arr(S),
                  _lib.f_arr(np.asarray(shifts, np.float32).ravel()), int(d == "forward"), out.vp)
        return out
