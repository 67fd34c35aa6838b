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
