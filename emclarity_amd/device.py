"""Device context and device arrays on top of the C ABI.

``DeviceArray`` is this package's stand-in for MATLAB's ``gpuArray``: a
device buffer plus a numpy shape and dtype. Shapes follow numpy C order, so a
volume of MATLAB size [nx, ny, nz] has shape (nz, ny, nx) and the same linear
element order. ``get()`` is MATLAB's ``gather``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class Context:
    """One HIP device + stream (emc_ctx)."""

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check("emc_ctx_create", lib.emc_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        _lib.call("emc_ctx_synchronize", self.handle)

    def wait_for(self, other: "Context"):
        """Later work on this context's stream waits for the work queued on
        ``other``'s stream so far (no host wait)."""
        _lib.call("emc_ctx_wait", self.handle, other.handle)

    def close(self):
        if self.handle:
            _lib.call("emc_ctx_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_DEFAULT: Context | None = None


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


class DeviceArray:
    """A device buffer with a numpy shape/dtype (owned unless ``owner`` is set)."""

    def __init__(self, shape, dtype, ctx: Context | None = None, _ptr=None, owner=None):
        self.ctx = ctx or default_context()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._owner = owner
        if _ptr is None:
            p = C.c_void_p()
            _lib.call("emc_malloc", self.ctx.handle, C.c_size_t(max(self.nbytes, 1)), C.byref(p))
            self.ptr = p.value
            self._owned = True
        else:
            self.ptr = _ptr
            self._owned = False

    # construction -------------------------------------------------------
    @classmethod
    def empty(cls, shape, dtype=np.float32, ctx=None):
        return cls(shape, dtype, ctx)

    @classmethod
    def zeros(cls, shape, dtype=np.float32, ctx=None):
        a = cls(shape, dtype, ctx)
        _lib.call("emc_memset_zero", a.ctx.handle, C.c_void_p(a.ptr), C.c_size_t(a.nbytes))
        return a

    @classmethod
    def from_numpy(cls, arr, dtype=None, ctx=None, wait=True):
        """A device copy of ``arr``. wait=False: staged through the context's
        pinned ring (emc_memcpy_h2d_async), ordered on its stream without
        waiting for the work already queued there; use the array on that
        context's stream only."""
        arr = np.ascontiguousarray(arr, dtype=dtype or arr.dtype)
        a = cls(arr.shape, arr.dtype, ctx)
        _lib.call("emc_memcpy_h2d" if wait else "emc_memcpy_h2d_async", a.ctx.handle, C.c_void_p(a.ptr),
                  arr.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))
        return a

    # access --------------------------------------------------------------
    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _lib.call("emc_memcpy_d2h", self.ctx.handle, out.ctypes.data_as(C.c_void_p),
                  C.c_void_p(self.ptr), C.c_size_t(self.nbytes))
        return out

    def copy(self) -> "DeviceArray":
        b = DeviceArray(self.shape, self.dtype, self.ctx)
        _lib.call("emc_memcpy_d2d", self.ctx.handle, C.c_void_p(b.ptr), C.c_void_p(self.ptr),
                  C.c_size_t(self.nbytes))
        return b

    def view(self, shape, dtype, offset_bytes=0) -> "DeviceArray":
        return DeviceArray(shape, dtype, self.ctx, _ptr=self.ptr + offset_bytes, owner=self)

    @property
    def vp(self):
        return C.c_void_p(self.ptr)

    @property
    def dims(self):
        """MATLAB size [nx, ny, nz] of a (nz, ny, nx) volume."""
        s = self.shape
        if len(s) == 3:
            return (s[2], s[1], s[0])
        if len(s) == 2:
            return (s[1], s[0], 1)
        return (s[0], 1, 1)

    def free(self):
        if self._owned and self.ptr and self.ctx.handle:
            _lib.call("emc_free", self.ctx.handle, C.c_void_p(self.ptr))
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def as_device(x, dtype=np.float32, ctx=None) -> DeviceArray:
    """numpy -> DeviceArray (copied); DeviceArray passes through."""
    if isinstance(x, DeviceArray):
        return x
    return DeviceArray.from_numpy(np.asarray(x), dtype=dtype, ctx=ctx)
