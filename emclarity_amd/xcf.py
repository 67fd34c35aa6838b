"""BH_multi_xcf_Translational / BH_multi_xcf_Rotational on the device."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray, as_device
from .fourier import FFTPlan


def BH_multi_xcf_Translational(rotPART_FT, REF_FT, peakMask, PEAK_COM, plan: FFTPlan | None = None,
                               return_max=False):
    """statistics/BH_multi_xcf_Translational.m: full-grid complex inputs
    (shape (nz,ny,nx) complex64), peakMask centred FP32 or None/'';
    returns the 1x3 shift (x,y,z). Raises EmcError(EMC_ERR_RANGE) when the
    COM box leaves the map, as the reference errors."""
    A = as_device(rotPART_FT, np.complex64)
    B = as_device(REF_FT, np.complex64, A.ctx)
    dims = A.dims
    plan = plan or FFTPlan(dims, 1, A.ctx)
    mask = None
    if peakMask is not None and not isinstance(peakMask, str):
        mask = as_device(peakMask, np.float32, A.ctx)
    work = DeviceArray((2,) + A.shape, np.complex64, A.ctx)
    peak = (C.c_double * 3)()
    mx = C.c_double()
    com = [int(c) for c in np.broadcast_to(np.asarray(PEAK_COM), (3,))]
    _lib.call("emc_xcf_translational", A.ctx.handle, plan.handle, A.vp, B.vp, _lib.i3(dims),
              mask.vp if mask is not None else None, _lib.i3(com), work.vp, peak, C.byref(mx))
    p = np.array(peak[:])
    return (p, mx.value) if return_max else p


def BH_multi_xcf_Rotational(rotPART_FT, REF_FT, wdgMask, refWdg, wCCC=0):
    """statistics/BH_multi_xcf_Rotational.m, un-shelled form (wCCC not a cell).
    Weights: full-grid FP32 arrays or scalars (scalar 1 / None = no weight)."""
    if isinstance(wCCC, (list, tuple)):
        raise NotImplementedError("shell-weighted CCC (wCCC cell) is not implemented")
    P = as_device(rotPART_FT, np.complex64)
    R = as_device(REF_FT, np.complex64, P.ctx)

    def wt(w):
        if w is None or (np.isscalar(w) and float(w) == 1.0):
            return None
        if np.isscalar(w):
            return DeviceArray.from_numpy(np.full(P.shape, w, np.float32), ctx=P.ctx)
        return as_device(w, np.float32, P.ctx)

    wp, wr = wt(wdgMask), wt(refWdg)
    cc = C.c_double()
    _lib.call("emc_xcf_rotational", P.ctx.handle, P.vp, R.vp, wp.vp if wp is not None else None,
              wr.vp if wr is not None else None, _lib.i3(P.dims), C.byref(cc))
    return cc.value
