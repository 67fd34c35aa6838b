"""Autotuner of the batched search's launch shapes.

Seven parameters shape the batched path's hot launches (emc_tune_set):
  search_tile  kx extent of the (kx, ky) column patches of the search
               schedule (ky extent 256 / search_tile): the shape of the
               reference-table region the units of one pass read together
  search_zsplit  kz segments per column, each its own pass: a pass reads
               1/search_zsplit of its patch's table region (more passes,
               each with its own y/x transform of the partial sums)
  search_passes  passes of that schedule per launch: they share the launch's
               last, partly filled round of workgroups
  grad_kernel  adjoint evaluation order: brick-major (1) or support-list (0)
  grad_brick   brick side of the brick-major adjoint (16 or 32)
  grad_waves   waves per SIMD the 32-brick adjoint's registers are
               allocated for (4 or 5)
  grad_blocks  workgroups a support-list adjoint evaluation aims for: how
               finely each particle's support is split (fill against fold
               work)

``python -m emclarity_amd.autotune [--out PATH]`` times every candidate on
the device with HIP events on the library's stream, for each box size of the
alignment workload (64/96/128^3 with 1,000/600/400 particles, the 28
orientations of the 10-degree coarse grid, scored over 0.6 x the support
radius as the bench's search is) and both table formats, and writes the fastest to a JSON
table. The table measured on MI355X ships as ``tuning/mi355x.json``;
BatchAligner applies its entries, and falls back to the library defaults
(16 x 16 patches, 2 passes per launch, 4096) for box sizes it does not
list.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib

CANDIDATES = {"search_tile": [4, 8, 16, 32, 64], "search_zsplit": [1, 2, 3, 4], "search_passes": [1, 2, 3, 4],
              "grad_kernel": [0, 1], "grad_brick": [16, 32], "grad_waves": [4, 5],
              "grad_blocks": [2048, 4096, 8192, 16384, 32768, 65536]}
TABLE = Path(__file__).resolve().parent / "tuning" / "mi355x.json"
_CACHE: dict | None = None


def load_table(path=None) -> dict:
    global _CACHE
    if path is None and _CACHE is not None:
        return _CACHE
    p = Path(path) if path else TABLE
    try:
        with open(p) as f:
            t = json.load(f)
    except OSError:
        t = {}
    if path is None:
        _CACHE = t
    return t


def params_for(n: int, table: str, tuning: dict | None = None) -> dict:
    """The tuned parameters for box size n and a table format ({} if none)."""
    t = tuning if tuning is not None else load_table()
    return dict(t.get("params", {}).get(table, {}).get(str(int(n)), {}))


def apply(ctx, n: int, params: dict):
    for name, value in params.items():
        _lib.call("emc_tune_set", ctx.handle, name.encode(), int(n), int(value))


def _timer(ctx, names):
    total = 0.0
    for name in names:
        ms, cnt = C.c_double(), C.c_int64()
        _lib.call("emc_timing_read", ctx.handle, name.encode(), C.byref(ms), C.byref(cnt))
        total += ms.value
    return total


def _timed(ctx, fn, names, reps):
    fn()                                      # warm-up: schedules, scratch
    ctx.synchronize()
    _lib.call("emc_timing_reset", ctx.handle)
    _lib.call("emc_timing_enable", ctx.handle, 1)
    for _ in range(reps):
        fn()
    ctx.synchronize()
    ms = _timer(ctx, names) / reps
    _lib.call("emc_timing_enable", ctx.handle, 0)
    return ms


def tune(ctx, counts=((64, 1000), (96, 600), (128, 400)), tables=("c64", "f16"), reps=3, seed=7,
         search_frac=0.6, log=sys.stderr):
    """Time every candidate; returns the table (params + all timings, ms).
    Per box size ``count`` particles, each with its own random start
    orientation (the launches' work and table access pattern depend on the
    box size, the support and the rotations, not on the voxel values: noise
    volumes stand in for the particles)."""
    from .batched import BatchAligner
    from .device import DeviceArray
    from .geometry import eulerSearch
    from .workload import make_size_class, random_rotation
    grid = eulerSearch("C1", 15, 10, 15, 10)
    out = {"device": "gfx950",
           "workload": "batched alignment, 28 coarse orientations (10 deg) over 0.6 x the support radius, "
                       "refinement evaluations over the full support, distinct random orientations, "
                       + ", ".join(f"{c} x {n}^3" for n, c in counts),
           "candidates": CANDIDATES, "params": {}, "timings_ms": {}}
    rng = np.random.default_rng(seed)
    for n, count in counts:
        sc, _ = make_size_class(n)
        data = DeviceArray.from_numpy(rng.standard_normal((count, n, n, n), dtype=np.float32), ctx=ctx)
        views = [data.view((n, n, n), np.float32, 4 * n ** 3 * i) for i in range(count)]
        starts = np.stack([random_rotation(rng) for _ in range(count)])
        for table in tables:
            al = BatchAligner(n, sc.ref, sc.bandpass, table=table, ctx=ctx, tuning={})
            al.set_particles(views, [None] * count, starts)
            Rd = DeviceArray.from_numpy(np.ascontiguousarray(starts.reshape(count, 9)), ctx=ctx)
            sd = DeviceArray.zeros((count, 3), np.float64, ctx)
            _lib.call("emc_tune_set", ctx.handle, b"refine_graph", n, 0)   # eager: events per evaluation
            best, times = {}, {}

            def sweep(name, timer):
                times[name] = {}
                for v in CANDIDATES[name]:
                    apply(ctx, n, {name: v})
                    ms = timer()
                    times[name][str(v)] = round(ms, 4)
                    print(f"[autotune] n={n} {table} {name}={v}: {ms:.3f} ms", file=log, flush=True)
                best[name] = int(min(CANDIDATES[name], key=lambda v: times[name][str(v)]))
                apply(ctx, n, {name: best[name]})
                return times[name][str(best[name])]

            def t_search():
                return _timed(ctx, lambda: al.search(grid, search_frac * al.support),
                              ("align_fused", "align_final"), reps)

            def t_eval():   # the refinement's evaluations (three per call)
                return _timed(ctx, lambda: al.refine_device(Rd, sd, iters=2), ("align_eval_grad",), reps)

            for name in ("search_tile", "search_zsplit", "search_passes"):
                sweep(name, t_search)
            # each adjoint form with its own knobs: the list form's workgroup
            # target, the brick form's side and waves per SIMD; then the faster
            apply(ctx, n, {"grad_kernel": 0})
            t_list = sweep("grad_blocks", t_eval)
            apply(ctx, n, {"grad_kernel": 1, "grad_waves": 4})
            sweep("grad_brick", t_eval)
            t_brick = sweep("grad_waves", t_eval)
            times["grad_kernel"] = {"0": t_list, "1": t_brick}
            best["grad_kernel"] = 0 if t_list <= t_brick else 1
            apply(ctx, n, {"grad_kernel": best["grad_kernel"]})
            print(f"[autotune] n={n} {table} grad_kernel={best['grad_kernel']} (list {t_list:.3f} ms, "
                  f"brick {t_brick:.3f} ms)", file=log, flush=True)
            out["params"].setdefault(table, {})[str(n)] = best
            out["timings_ms"].setdefault(table, {})[str(n)] = times
            del al
        del views, data
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="time the batched path's launch shapes and write the table")
    ap.add_argument("--out", default=str(TABLE))
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args(argv)
    from .device import default_context
    t0 = time.perf_counter()
    res = tune(default_context(), reps=a.reps)
    res["seconds"] = round(time.perf_counter() - t0, 1)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res["params"]))


if __name__ == "__main__":
    main()
