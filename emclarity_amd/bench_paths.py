"""Throughput of the paths around the alignment search, at the sizes of the
batched-alignment and template-search workloads:

  extract  200 windows of 128^3 cut from a 512 x 512 x 256 int16 tomogram
  average  100 particles of 128^3 with 128^3 wedges into the half-set sums
  peaks    greedy peak extraction on a 512 x 512 x 256 score map, 150 peaks
  tsprep   pre-processing of one 480^3 template-search chunk
  tsearch  one 480^3 chunk against a 48^3 template, 32 orientations

    python -m emclarity_amd.bench_paths [--path all|extract|average|peaks|tsprep] [--reps N]

Prints one JSON line per path: wall time per call and, where the library
times the dominant kernel with HIP events on its stream, that kernel's
average duration, its algorithmic bytes per launch and the rate.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _lib
from .device import DeviceArray, default_context


def _ktime(ctx, name):
    ms, cnt = C.c_double(), C.c_int64()
    _lib.call("emc_timing_read", ctx.handle, name.encode(), C.byref(ms), C.byref(cnt))
    return ms.value, cnt.value


def _timed(ctx, fn, reps):
    fn()                                   # warm-up: allocations, plans
    ctx.synchronize()
    _lib.call("emc_timing_reset", ctx.handle)
    _lib.call("emc_timing_enable", ctx.handle, 1)
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / reps
    _lib.call("emc_timing_enable", ctx.handle, 0)
    return dt


def _line(path, wall, extra):
    d = {"path": path, "ms_per_call": round(wall * 1e3, 3)}
    d.update(extra)
    print(json.dumps(d), flush=True)


def bench_extract(ctx, reps):
    from .subtomo import ResidentTomogram, extract_subtomos
    rng = np.random.default_rng(1)
    tomo = ResidentTomogram.from_numpy(rng.integers(-3000, 3000, (256, 512, 512), dtype=np.int16), ctx)
    n, w = 200, 128
    c = np.stack([rng.uniform(70, 440, n), rng.uniform(70, 440, n), rng.uniform(70, 185, n)], 1)
    out = DeviceArray((n, w, w, w), np.float32, ctx)
    wall = _timed(ctx, lambda: extract_subtomos(tomo, c, (w, w, w), 20.0, out=out), reps)
    ms, cnt = _ktime(ctx, "extract_subtomos")
    bytes_ = n * w ** 3 * (2 + 4)           # int16 read + float32 written per window voxel
    k = ms / max(cnt, 1)
    _line("extract", wall, {"kernel": "extract_kernel<1>", "kernel_ms": round(k, 4), "bytes": bytes_,
                            "GBps": round(bytes_ / (k * 1e-3) / 1e9, 1), "windows": n, "window": w})


def bench_average(ctx, reps):
    from .average import HalfSetSums, average_accumulate
    from .workload import random_rotation
    rng = np.random.default_rng(2)
    n, w = 100, 128
    wins = DeviceArray.from_numpy(rng.standard_normal((n, w, w, w), dtype=np.float32), ctx=ctx)
    wedges = DeviceArray.from_numpy(rng.uniform(0, 1, (n, w, w, w)).astype(np.float32), ctx=ctx)
    mats = np.stack([random_rotation(rng).reshape(-1, order="F") for _ in range(n)])
    sh = rng.uniform(-0.5, 0.5, (n, 3))
    half = 1 + (np.arange(n) % 2)
    z, y, x = np.meshgrid(*(np.arange(w) - w // 2,) * 3, indexing="ij")
    mask = np.clip((w / 2 - 6 - np.sqrt(x * x + y * y + z * z)) / 3.0, 0, 1).astype(np.float32)
    sums = HalfSetSums(w, w, ctx)
    wall = _timed(ctx, lambda: average_accumulate(sums, wins, mats, sh, half, np.ones(n), mask, wedges), reps)
    rot_ms, rot_cnt = _ktime(ctx, "avg_rotate")
    acc_ms, acc_cnt = _ktime(ctx, "avg_accumulate")
    v = w ** 3
    rot_bytes = n * v * 8                   # window read once + rotated copy written
    acc_bytes = n * v * 4 + v * 4 * 3       # rotated copies read + mask + one sum read and written
    _line("average", wall, {
        "kernel": "avg_rotate_kernel", "kernel_ms": round(rot_ms / max(rot_cnt, 1), 4), "bytes": rot_bytes,
        "GBps": round(rot_bytes / (rot_ms / max(rot_cnt, 1) * 1e-3) / 1e9, 1),
        "accumulate_ms": round(acc_ms / max(acc_cnt, 1), 4), "accumulate_bytes": acc_bytes,
        "accumulate_GBps": round(acc_bytes / (acc_ms / max(acc_cnt, 1) * 1e-3) / 1e9, 1),
        "particles": n, "window": w})


def bench_peaks(ctx, reps):
    from .masking import BH_mask3d
    rng = np.random.default_rng(3)
    nz, ny, nx = 256, 512, 512
    rm = 4 + 7
    mag = (0.05 * rng.standard_normal((nz, ny, nx))).astype(np.float32)
    ang = rng.integers(1, 4600, (nz, ny, nx)).astype(np.int32)
    pts = []
    while len(pts) < 150:
        p = np.array([rng.integers(20, nx - 20), rng.integers(20, ny - 20), rng.integers(20, nz - 20)])
        if all(np.max(np.abs(p - q)) > 16 for q in pts):
            pts.append(p)
    for (x, y, z) in pts:
        mag[z - 1:z + 2, y - 1:y + 2, x - 1:x + 2] += rng.uniform(2, 5)
    pm, pa = np.pad(mag, rm), np.pad(ang, rm)
    grid = np.stack([np.arange(4600) % 360.0, np.full(4600, 15.0), np.zeros(4600)], 1)
    mask = BH_mask3d("sphere", (2 * rm + 1,) * 3, (4.0, 4.0, 4.0), [0, 0, 0])
    d_ang = DeviceArray.from_numpy(pa, ctx=ctx)
    pd = C.POINTER(C.c_double)
    out = np.zeros((300, 7))
    nf = C.c_int()
    g = np.ascontiguousarray(grid)
    d_mag = DeviceArray(pm.shape, np.float32, ctx)

    def run():
        _lib.call("emc_memcpy_h2d", ctx.handle, d_mag.vp, pm.ctypes.data_as(C.c_void_p), C.c_size_t(pm.nbytes))
        _lib.call("emc_tsearch_peaks", ctx.handle, d_mag.vp, d_ang.vp, _lib.i3(d_mag.dims), mask.vp, rm,
                  g.ctypes.data_as(pd), 4600, 0, 0.98, 1.0, 300, 10000, out.ctypes.data_as(pd), C.byref(nf))

    wall = _timed(ctx, run, reps)
    _line("peaks", wall, {"peaks": int(nf.value), "map": [nx, ny, nz], "ms_per_peak": round(wall * 1e3 / max(nf.value, 1), 4),
                          "note": "wall time includes the 270 MB host-to-device copy of the map"})


def bench_tsprep(ctx, reps):
    rng = np.random.default_rng(4)
    s = 480
    chunk = DeviceArray.from_numpy(rng.standard_normal((s, s, s), dtype=np.float32), ctx=ctx)
    out = DeviceArray((s, s, s), np.float32, ctx)
    bp = np.array([1e-3, 600.0, 28.0, 10.0])
    total = C.c_double()
    wall = _timed(ctx, lambda: _lib.call("emc_tsearch_preprocess_chunk", ctx.handle, chunk.vp, _lib.i3((s, s, s)),
                                         bp.ctypes.data_as(C.POINTER(C.c_double)), 0, 8, -1.0, None, out.vp,
                                         C.byref(total)), reps)
    _line("tsprep", wall, {"chunk": s, "GB_in_out": round(2 * 4 * s ** 3 / 1e9, 3)})


def bench_tsearch(ctx, reps):
    from .geometry import eulerSearch
    rng = np.random.default_rng(5)
    s, t, n_orient = 480, 48, 32
    chunk = DeviceArray.from_numpy(rng.standard_normal((s, s, s), dtype=np.float32), ctx=ctx)
    tmpl = DeviceArray.from_numpy(rng.standard_normal((t, t, t), dtype=np.float32), ctx=ctx)
    pts = np.ascontiguousarray(eulerSearch("C1", 180, 15, 180, 15)[:n_orient])
    vlo, vd = 24, s - 48
    peak = DeviceArray.zeros((vd, vd, vd), np.float32, ctx)
    ang = DeviceArray.zeros((vd, vd, vd), np.int32, ctx)
    pl = (s - t) // 2

    def run():
        _lib.call("emc_template_search_chunk", ctx.handle, chunk.vp, _lib.i3((s, s, s)), tmpl.vp, _lib.i3((t, t, t)),
                  _lib.i3((pl, pl, pl)), _lib.i3((vlo, vlo, vlo)), _lib.i3((vd, vd, vd)),
                  pts.ctypes.data_as(C.POINTER(C.c_double)), n_orient, 1, 1, peak.vp, ang.vp)

    wall = _timed(ctx, run, reps)
    hc = (s // 2 + 1) * s * s
    b_ts = 8 * t ** 3 + 12 * s ** 3 + 40 * hc + 24 * vd ** 3      # per (chunk, orientation)
    per = wall / n_orient
    _line("tsearch", wall, {"chunk": s, "template": t, "orientations": n_orient,
                            "ms_per_orientation": round(per * 1e3, 4), "bytes_per_orientation": b_ts,
                            "GBps": round(b_ts / per / 1e9, 1)})


PATHS = {"extract": bench_extract, "average": bench_average, "peaks": bench_peaks, "tsprep": bench_tsprep,
         "tsearch": bench_tsearch}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--path", default="all", choices=["all", *PATHS])
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args(argv)
    ctx = default_context()
    for name in (PATHS if a.path == "all" else [a.path]):
        print(f"[bench_paths] {name}", file=sys.stderr, flush=True)
        PATHS[name](ctx, a.reps)


if __name__ == "__main__":
    main()
