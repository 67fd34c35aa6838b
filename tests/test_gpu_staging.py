"""Host-to-device uploads through the context's pinned staging ring
(emc_memcpy_h2d_async): every upload lands intact when the ring wraps
several times with uploads still in flight, when the host buffer is
overwritten right after the call returns, and behind queued device work;
uploads above a quarter of the ring take the waiting path."""
import ctypes as C

import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def test_async_uploads_survive_ring_wrap_and_host_reuse(ctx):
    rng = np.random.default_rng(3)
    sizes = [int(s) for s in rng.integers(1, 3 << 20, 60)]      # up to 3 MiB each: ~90 MiB through 32 MiB
    host = np.empty(max(sizes) // 4 + 1, np.float32)
    expected, dev = [], []
    for i, sz in enumerate(sizes):
        n = sz // 4 + 1
        vals = rng.standard_normal(n).astype(np.float32)
        host[:n] = vals                                          # one host buffer, rewritten every call
        d = DeviceArray((n,), np.float32, ctx)
        _lib.call("emc_memcpy_h2d_async", ctx.handle, d.vp, host.ctypes.data_as(C.c_void_p), C.c_size_t(4 * n))
        host[:n] = -1.0                                          # the call returned: src may be reused
        expected.append(vals)
        dev.append(d)
    for vals, d in zip(expected, dev):
        assert np.array_equal(d.get(), vals)


def test_async_upload_is_ordered_behind_queued_work(ctx):
    # a large device copy queued first, then a small upload into its source:
    # the copy must still see the old values (stream order)
    n = 64 << 20                                                  # 256 MiB
    src = DeviceArray.from_numpy(np.full(n, 7.0, np.float32), ctx=ctx)
    dst = DeviceArray((n,), np.float32, ctx)
    _lib.call("emc_memcpy_d2d", ctx.handle, dst.vp, src.vp, C.c_size_t(4 * n))
    small = np.full(1024, 3.0, np.float32)
    _lib.call("emc_memcpy_h2d_async", ctx.handle, src.vp, small.ctypes.data_as(C.c_void_p), C.c_size_t(4096))
    head = dst.get()[:1024]
    assert np.all(head == 7.0)
    assert np.all(src.get()[:1024] == 3.0)
    big = np.arange(4 << 20, dtype=np.float32)                   # 16 MiB: above a quarter of the ring
    d = DeviceArray((big.size,), np.float32, ctx)
    _lib.call("emc_memcpy_h2d_async", ctx.handle, d.vp, big.ctypes.data_as(C.c_void_p), C.c_size_t(big.nbytes))
    assert np.array_equal(d.get(), big)
