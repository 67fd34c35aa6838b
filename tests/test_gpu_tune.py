"""Launch-shape parameters on the device: names, ranges, the per-box
override, and that the search's results do not depend on the patch tile or
the kz segments per column beyond FP32 summation order (same best
orientation, scores within 1e-5)."""
import ctypes as C

import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.batched import BatchAligner
from emclarity_amd.geometry import eulerSearch
from emclarity_amd.workload import make_particle, make_size_class

pytestmark = pytest.mark.gpu


def _get(ctx, name, box):
    v = C.c_int()
    _lib.call("emc_tune_get", ctx.handle, name.encode(), box, C.byref(v))
    return v.value


def test_tune_parameters(ctx):
    from emclarity_amd.device import Context
    ctx = Context(0)                # fresh: the shared context carries the shipped tuning
    assert _get(ctx, "search_tile", 64) == 16 and _get(ctx, "grad_blocks", 64) == 4096
    _lib.call("emc_tune_set", ctx.handle, b"search_tile", 0, 12)
    _lib.call("emc_tune_set", ctx.handle, b"search_tile", 96, 24)
    assert _get(ctx, "search_tile", 64) == 12 and _get(ctx, "search_tile", 96) == 24
    for bad in ((b"search_tile", 0, 1), (b"search_tile", 0, 129), (b"grad_blocks", 0, 100), (b"nope", 0, 8)):
        with pytest.raises(_lib.EmcError):
            _lib.call("emc_tune_set", ctx.handle, *bad)
    ctx.close()


def test_search_independent_of_tile(ctx):
    n = 64
    rng = np.random.default_rng(3)
    sc, blobs = make_size_class(n)
    parts = [make_particle(rng, sc, blobs) for _ in range(4)]
    grid = eulerSearch("C1", 15, 6, 15, 6)
    out = []
    # patch tiles, and kz segments per column (search_zsplit: a column's kz
    # range split over several passes)
    for tile, zs in ((16, 1), (4, 1), (64, 1), (16, 2), (16, 4), (8, 3)):
        al = BatchAligner(n, sc.ref, sc.bandpass, ctx=ctx,
                          tuning={"params": {"c64": {"64": {"search_tile": tile, "search_zsplit": zs}}}})
        al.set_particles([p.data for p in parts], [None] * 4, [p.start for p in parts])
        out.append(al.search(grid))
    _lib.call("emc_tune_set", ctx.handle, b"search_tile", 64, 16)
    _lib.call("emc_tune_set", ctx.handle, b"search_zsplit", 64, 1)
    for score, shift in out[1:]:
        assert np.array_equal(np.argmax(score, 1), np.argmax(out[0][0], 1))
        assert np.max(np.abs(score - out[0][0])) < 1e-5
        assert np.max(np.abs(shift - out[0][1])) < 1e-3


def test_adjoint_independent_of_kernel_form(ctx):
    """The adjoint's score and gradient through every kernel form: the
    support-list kernel over chunks of ~255 to ~1,050 entries (ragged
    against 256 threads), and the brick-major kernel with 16 and 32 bricks
    (4 and 5 waves per SIMD); all agree to FP32 summation order, and two
    particles against the FP64 oracle."""
    from emclarity_amd.batched import euler_derivatives
    from oracle import emc_oracle as O
    n, P = 64, 16
    rng = np.random.default_rng(8)
    sc, blobs = make_size_class(n)
    parts = [make_particle(rng, sc, blobs) for _ in range(P)]
    R, dR = zip(*[euler_derivatives(p.start, *rng.uniform(-8.0, 8.0, 3)) for p in parts])
    R, dR = np.stack(R), np.stack([np.stack(d) for d in dR])
    s = rng.uniform(-2.0, 2.0, (P, 3))
    forms = [(0, 32, 4, 256), (0, 32, 4, 1024), (0, 32, 4, 4096), (1, 32, 4, 4096), (1, 32, 5, 4096),
             (1, 16, 4, 4096)]
    res = []
    for kind, brick, waves, blocks in forms:
        t = {"grad_kernel": kind, "grad_brick": brick, "grad_waves": waves, "grad_blocks": blocks}
        al = BatchAligner(n, sc.ref, sc.bandpass, ctx=ctx, tuning={"params": {"c64": {"64": t}}})
        al.set_particles([p.data for p in parts], [None] * P, [p.start for p in parts])
        res.append(al.eval_grad(R, dR, s))
    for name, dflt in (("grad_kernel", 1), ("grad_brick", 32), ("grad_waves", 4), ("grad_blocks", 4096)):
        _lib.call("emc_tune_set", ctx.handle, name.encode(), 64, dflt)
    cc0, g0 = res[0]
    for cc, g in res[1:]:
        assert np.max(np.abs(cc - cc0)) < 1e-5 * np.max(np.abs(cc0))
        assert np.max(np.abs(g - g0)) < 1e-4 * np.max(np.abs(g0))
    F2 = O.batch_reference_spectrum(sc.ref, "c64")
    for i in (0, P - 1):
        bpp = O.batch_prepare(parts[i].data, sc.bandpass, None, al.support)
        ocs, og = O.batch_eval(F2, n, bpp, R[i], s[i], dR[i])
        for cc, g in res:
            assert abs(cc[i] - ocs) < 1e-4 * abs(ocs)
            assert np.max(np.abs(g[i] - og)) < 2e-3 * np.max(np.abs(og))
