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


class ShellPlan:
    """The shells of a wCCC cell resident on the device: wCCC{1} (abs flag),
    wCCC{2} (the constant w of f(c) = (c^2 / (c + w^-1.5))^3) and the voxel
    lists wCCC{3..end}, 1-based linear indices into the full grid in MATLAB
    (column-major) order, as BH_multi_cRef's calc_shells makes them
    (statistics/BH_multi_cRef.m:168-216). The reference moves the cell to the
    GPU once per reference (alignment/BH_alignRaw3d_v2.m:832-842); make one
    ShellPlan per cell likewise and pass it as wCCC."""

    def __init__(self, wCCC, ctx=None):
        from .device import default_context
        self.ctx = ctx or default_context()
        cell = list(wCCC)
        if len(cell) < 3:
            raise ValueError("a wCCC cell holds the abs flag, w and at least one shell")
        self.abs_flag = int(bool(np.asarray(cell[0]).ravel()[0]))
        self.w = float(np.asarray(cell[1]).ravel()[0])
        shells = [np.asarray(c).ravel().astype(np.int64) for c in cell[2:]]
        for sh in shells:
            if sh.size and (sh.min() < 1 or sh.max() > 0xFFFFFFFF):
                raise ValueError("shell indices are 1-based uint32 linear indices")
        self.n_shells = len(shells)
        off = np.zeros(self.n_shells + 1, np.int64)
        off[1:] = np.cumsum([sh.size for sh in shells])
        idx = np.ascontiguousarray(np.concatenate(shells).astype(np.uint32)) if off[-1] else np.zeros(1, np.uint32)
        h = C.c_void_p()
        _lib.call("emc_shell_plan_create", self.ctx.handle, self.abs_flag, self.w, self.n_shells,
                  off.ctypes.data_as(C.POINTER(C.c_int64)), idx.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(h))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            _lib.call("emc_shell_plan_destroy", self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def BH_multi_xcf_Rotational(rotPART_FT, REF_FT, wdgMask, refWdg, wCCC=0):
    """statistics/BH_multi_xcf_Rotational.m. Weights: full-grid FP32 arrays
    or scalars (scalar 1 / None = no weight); a weight whose first element is
    NaN counts as 1 (:34-40). wCCC not a cell: the plain CCC (:71-78); a
    wCCC cell (list/tuple) or a ShellPlan: the shell-weighted sum (:50-70).
    Volumes are (nz, ny, nx) arrays, so MATLAB's 1-based column-major linear
    index i names flat element i-1."""
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
    if isinstance(wCCC, (list, tuple, ShellPlan)):
        plan = wCCC if isinstance(wCCC, ShellPlan) else ShellPlan(wCCC, P.ctx)
        _lib.call("emc_xcf_rotational_shells", P.ctx.handle, plan.handle, P.vp, R.vp,
                  wp.vp if wp is not None else None, wr.vp if wr is not None else None, _lib.i3(P.dims),
                  C.byref(cc))
        return cc.value
    _lib.call("emc_xcf_rotational", P.ctx.handle, P.vp, R.vp, wp.vp if wp is not None else None,
              wr.vp if wr is not None else None, _lib.i3(P.dims), C.byref(cc))
    return cc.value
