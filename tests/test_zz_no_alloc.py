"""emc_align_eval_grad_noalloc (the allocation-free adjoint evaluation)
against emc_align_eval_grad through the C ABI: bit-identical scores and
gradients on two calls in a row (128^3 particles, then a smaller input:
32^3 and fewer particles), the same pair again (support list reused from
the static array), and the size limit rejected with a clear error."""
import ctypes as C

import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.batched import BatchAligner, euler_derivatives
from emclarity_amd.device import DeviceArray
from emclarity_amd.workload import random_rotation

pytestmark = pytest.mark.gpu


def _case(ctx, n, count, seed, table="c64"):
    rng = np.random.default_rng(seed)
    ax = np.arange(n) - n // 2
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    ref = np.exp(-((x - 2.0) ** 2 + (y + 1.0) ** 2 + z ** 2) / (2 * (0.12 * n) ** 2)).astype(np.float32)
    ref += 0.3 * np.exp(-((x + 0.2 * n) ** 2 + y ** 2 + (z - 0.1 * n) ** 2) / (2 * (0.06 * n) ** 2)).astype(np.float32)
    f = np.fft.fftfreq(n)
    kz, ky, kx = np.meshgrid(f, f, f, indexing="ij")
    bp = np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) / (2 * 0.15 ** 2)).astype(np.float32)
    al = BatchAligner(n, ref, bp, support_radius=0.35 * n, peak_radius=6, table=table, ctx=ctx)
    parts = [rng.standard_normal((n, n, n)).astype(np.float32) + ref for _ in range(count)]
    starts = [random_rotation(rng) for _ in range(count)]
    al.set_particles(parts, [None] * count, starts)
    R, dR = zip(*[euler_derivatives(s, *rng.uniform(-5, 5, 3)) for s in starts])
    h = np.zeros((count, 36), np.float32)
    h[:, :9] = 2.0 * np.transpose(np.stack(R), (0, 2, 1)).reshape(count, 9)
    h[:, 9:] = 2.0 * np.transpose(np.stack([np.stack(d) for d in dR]), (0, 1, 3, 2)).reshape(count, 27)
    d_m = DeviceArray.from_numpy(h, ctx=ctx)
    d_s = DeviceArray.from_numpy(rng.uniform(-2, 2, (count, 3)).astype(np.float32), ctx=ctx)
    return al, d_m, d_s


def _eval(ctx, name, al, d_m, d_s):
    out = DeviceArray((al.n_part, 7), np.float64, ctx)
    _lib.call(name, ctx.handle, al.n, al.n_part, al._fmt, al.F2.vp, al.Q.vp, al.N.vp, al.A.vp, d_m.vp, d_s.vp,
              al.support, out.vp)
    return out.get()


def test_noalloc_eval_grad_matches_the_allocating_entry(ctx):
    big = _case(ctx, 128, 12, 1)
    small = _case(ctx, 32, 3, 2, table="f16")
    for case in (big, small, small, big):
        al = case[0]
        al._apply_tuning()
        a = _eval(ctx, "emc_align_eval_grad", *case)
        b = _eval(ctx, "emc_align_eval_grad_noalloc", *case)
        assert np.all(np.isfinite(a)) and np.array_equal(a, b), (al.n, np.max(np.abs(a - b)))


def test_noalloc_eval_grad_rejects_inputs_beyond_its_scratch(ctx):
    al, d_m, d_s = _case(ctx, 32, 2, 3)
    out = DeviceArray((2, 7), np.float64, ctx)
    # 200,000 particles of 128^3 need far more partial sums than 16 MiB hold;
    # the check runs before any launch, so the small buffers are never read
    with pytest.raises(_lib.EmcError, match="scratch"):
        _lib.call("emc_align_eval_grad_noalloc", ctx.handle, 128, 200000, 0, al.F2.vp, al.Q.vp, al.N.vp, al.A.vp,
                  d_m.vp, d_s.vp, 60.0, out.vp)
    _eval(ctx, "emc_align_eval_grad_noalloc", al, d_m, d_s)   # the library state is intact
