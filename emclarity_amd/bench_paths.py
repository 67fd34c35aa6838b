"""Throughput of the paths around the alignment search, at the sizes of the
batched-alignment and template-search workloads:

  extract  200 windows of 128^3 cut from a 512 x 512 x 256 int16 tomogram
  average  100 particles of 128^3 with 128^3 wedges into the half-set sums
  peaks    greedy peak extraction on a 512 x 512 x 256 score map, 150 peaks
  tsprep   pre-processing of one 480^3 template-search chunk
  tsearch  one 480^3 chunk against a 48^3 template, 32 orientations
  adjoint  one adjoint evaluation of 1,000 / 600 / 400 particles of 64 / 96 / 128^3
  refine   the refinement's adjoint evaluations on the same size classes
  search   the 28-orientation coarse search of the same size classes
  copy     a 1 GiB device-to-device copy (byte-counter calibration)
  wedges   the batched wedge build (emc_weight_mask_batched) of 1,000 / 600 /
           400 particles of 64 / 96 / 128^3, each its own tilt range U[50, 70]
           deg in 3-degree steps and defocus U[1.5, 4] um
  plain    the drop-in per-particle grid search (emc_align_grid) of one
           particle of 64 / 96 / 128^3 over the 1,045-orientation local grid,
           orientations in blocks; the rotation kernel's rate, and the whole
           per-particle alignment (grid + refinement + final pass)
  pipeline template search of a 1024 x 1024 x 512 tomogram holding 600 copies
           of a 64^3 template (SNR 0.03, random orientations) over the
           7-degree grid as one whole-tomogram chunk, greedy extraction of
           the top 1,000 peaks, 96^3 windows under a soft sphere of radius
           26, batched adjoint refinement;
           recovery and pose errors against the inserted copies, and the
           plain per-particle grid search timed on a sample of the windows
  gradient 256 particles of 128^3 (+-60 deg wedge, own defocus, SNR 0.1):
           the adjoint against central differences, and gradient refinement
           from the 10-degree coarse peak (judged by the grid) against the
           dense 1-degree local grid (23,281 orientations), accuracy and time

    python -m emclarity_amd.bench_paths [--path all|extract|...|copy] [--reps N] [--table c64|f16]

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
    # with the optional 3^3 and 5^3 median filters first (med_filt 3, 5)
    wall = _timed(ctx, lambda: _lib.call("emc_tsearch_preprocess_chunk", ctx.handle, chunk.vp, _lib.i3((s, s, s)),
                                         bp.ctypes.data_as(C.POINTER(C.c_double)), 3, 8, -1.0, None, out.vp,
                                         C.byref(total)), reps)
    _line("tsprep_median3", wall, {"chunk": s, "GB_in_out": round(2 * 4 * s ** 3 / 1e9, 3)})
    wall = _timed(ctx, lambda: _lib.call("emc_tsearch_preprocess_chunk", ctx.handle, chunk.vp, _lib.i3((s, s, s)),
                                         bp.ctypes.data_as(C.POINTER(C.c_double)), 5, 8, -1.0, None, out.vp,
                                         C.byref(total)), reps)
    _line("tsprep_median5", wall, {"chunk": s, "GB_in_out": round(2 * 4 * s ** 3 / 1e9, 3)})


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

    hc = (s // 2 + 1) * s * s
    b_ts = 8 * t ** 3 + 12 * s ** 3 + 40 * hc + 24 * vd ** 3      # per (chunk, orientation)
    v = C.c_int()
    _lib.call("emc_tune_get", ctx.handle, b"tsearch_pruned", 0, C.byref(v))
    default = v.value
    for mode in (2, 1, 0):   # matrix-core z stage, vector z stage, padded-volume R2C
        _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, mode)
        wall = _timed(ctx, run, reps)
        per = wall / n_orient
        _line("tsearch", wall, {"chunk": s, "template": t, "orientations": n_orient, "tsearch_pruned": mode,
                                "ms_per_orientation": round(per * 1e3, 4), "bytes_per_orientation": b_ts,
                                "GBps": round(b_ts / per / 1e9, 1)})
    _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, default)


# ------------------------------------------------ batched-alignment kernels
# The alignment workload's size classes (box, particles). The kernels' work
# and memory access pattern depend on the box size, the support radius and
# the particles' rotations, not on the voxel values: noise particles with
# random start orientations stand in for the rendered ones.
CLASSES = ((64, 1000), (96, 600), (128, 400))


def _noise_class(ctx, n, count, table, seed):
    from .batched import BatchAligner
    from .workload import make_size_class, random_rotation
    rng = np.random.default_rng(seed)
    sc, _ = make_size_class(n)
    al = BatchAligner(n, sc.ref, sc.bandpass, peak_radius=6, table=table, ctx=ctx)
    data = DeviceArray.from_numpy(rng.standard_normal((count, n, n, n), dtype=np.float32), ctx=ctx)
    nv = n ** 3
    views = [data.view((n, n, n), np.float32, 4 * nv * i) for i in range(count)]
    starts = np.stack([random_rotation(rng) for _ in range(count)])
    al.set_particles(views, [None] * count, starts)
    del views, data
    return al, starts


def table_entries_in_ball(n, radius):
    """Entries of the centred reference table (x in [0, n+1], y, z in [-n, n])
    whose 2x2x2 footprints trilinear samples at |j| <= radius read."""
    r = radius + 2.0
    x = np.arange(0, n + 2)[:, None, None]
    y = np.arange(-n, n + 1)[None, :, None]
    z = np.arange(-n, n + 1)[None, None, :]
    return int(np.count_nonzero(x * x + y * y + z * z <= r * r))


def eval_grad_bytes(n, count, support, table):
    """Algorithmic bytes of one adjoint evaluation of ``count`` particles:
    each particle's Q (8 B) and N (4 B) at every support frequency, the
    support list once (8 B an entry), the table region the samples reach
    once (32 B an entry complex64, 16 B fp16). Returns (bytes, samples)."""
    from .batched import support_voxels
    ns = support_voxels(n, support) - 1                       # DC excluded
    entry = 32 if table == "c64" else 16
    tab = table_entries_in_ball(n, 2.0 * support) * entry
    return count * ns * 12 + ns * 8 + tab, count * ns


def bench_adjoint(ctx, reps, table="c64"):
    """One adjoint evaluation (score + 6-component gradient) of every
    particle of a size class at its pose, full support: the refinement's
    kernel (align_eval_grad_kernel), HIP events on the library stream."""
    for k, (n, count) in enumerate(CLASSES):
        al, starts = _noise_class(ctx, n, count, table, 60 + k)
        s = np.zeros((count, 3))
        wall = _timed(ctx, lambda: al._eval_local(starts, s), reps)
        ms, cnt = _ktime(ctx, "align_eval_grad")
        per = ms / max(cnt, 1)
        b, samples = eval_grad_bytes(n, count, al.support, table)
        _line("adjoint", wall, {"kernel": f"align_eval_grad_kernel<{table}>", "n": n, "particles": count,
                                "support_radius": round(al.support, 3), "samples": samples,
                                "kernel_ms": round(per, 4), "bytes": b, "GBps": round(b / (per * 1e-3) / 1e9, 1),
                                "hbm_frac": round(b / (per * 1e-3) / 8e12, 4),
                                "gsamples_per_s": round(samples / (per * 1e-3) / 1e9, 2)})
        del al


def bench_refine(ctx, reps, table="c64", iters=8):
    """The refinement's evaluations: emc_align_refine with eager launches
    (HIP events per adjoint kernel) over the full support from each
    particle's start pose; the kernel and brick side from the tuning table
    unless REFINE_KERNEL / REFINE_BRICK say otherwise."""
    import os
    kern = os.environ.get("REFINE_KERNEL")
    brick = os.environ.get("REFINE_BRICK")
    for k, (n, count) in enumerate(CLASSES):
        al, starts = _noise_class(ctx, n, count, table, 80 + k)
        al._apply_tuning()
        _lib.call("emc_tune_set", ctx.handle, b"refine_graph", n, 0)
        if kern is not None:
            _lib.call("emc_tune_set", ctx.handle, b"grad_kernel", n, int(kern))
        if brick is not None:
            _lib.call("emc_tune_set", ctx.handle, b"grad_brick", n, int(brick))
        R = DeviceArray.from_numpy(np.ascontiguousarray(starts.reshape(count, 9)), ctx=ctx)
        s = DeviceArray.zeros((count, 3), np.float64, ctx)
        wall = _timed(ctx, lambda: al.refine_device(R, s, iters=iters), reps)
        ms, cnt = _ktime(ctx, "align_eval_grad")
        per = ms / max(cnt, 1)
        b, samples = eval_grad_bytes(n, count, al.support, table)
        v = C.c_int()
        _lib.call("emc_tune_get", ctx.handle, b"grad_kernel", n, C.byref(v))
        _line("refine", wall, {"kernel": f"adjoint<{table}, {'brick' if v.value else 'list'}>", "n": n,
                               "particles": count, "evals_per_call": round(cnt / reps, 2), "kernel_ms": round(per, 4),
                               "bytes": b, "GBps": round(b / (per * 1e-3) / 1e9, 1),
                               "gsamples_per_s": round(samples / (per * 1e-3) / 1e9, 2)})
        del al


def bench_search(ctx, reps, table="c64", frac=0.6):
    """The coarse search of a size class: 28 orientations of the 10-degree
    local grid scored over 0.6 x the support radius (align_pass_kernel +
    align_final_kernel)."""
    from .batched import fused_samples_per_unit
    from .geometry import eulerSearch
    grid = eulerSearch("C1", 15, 10, 15, 10)
    for k, (n, count) in enumerate(CLASSES):
        al, _ = _noise_class(ctx, n, count, table, 70 + k)
        wall = _timed(ctx, lambda: al.search_device(grid, support_radius=frac * al.support), reps)
        ms, cnt = _ktime(ctx, "align_fused")
        fms, fcnt = _ktime(ctx, "align_final")
        units = count * len(grid)
        samples = fused_samples_per_unit(n, frac * al.support) * units
        _line("search", wall, {"kernel": f"align_pass_kernel<6, {table}>", "n": n, "particles": count,
                               "orientations": len(grid), "launches_per_call": round(cnt / reps, 2),
                               "pass_ms_per_call": round(ms / reps, 4), "final_ms_per_call": round(fms / reps, 4),
                               "gsamples_per_s": round(samples / (ms / reps * 1e-3) / 1e9, 2)})
        del al


def bench_plain(ctx, reps):
    """emc_align_grid over the local grid (both stages, 5 rotated volumes of
    n^3 per orientation: reference, reference weight, peak mask, particle
    and wedge), then align_particle as the bench's plain path runs it. The
    rotation kernel's bytes: one read of the input and one write of the
    output per rotated voxel (8 B; the particle's window-mask read is not
    counted)."""
    from . import synthetic as S
    from .align_raw import ParticleAligner, align_particle
    from .masking import BH_weightMaskMex
    from .workload import local_grid, make_particle, make_size_class
    grid = local_grid()
    v = _lib.C.c_int()
    for n in (64, 96, 128):
        rng = np.random.default_rng(500 + n)
        sc, blobs = make_size_class(n)
        p = make_particle(rng, sc, blobs)
        w = BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]], ctx=ctx).get()
        al = ParticleAligner(p.data, sc.ref, w, sc.bandpass, sc.peak_mask, vol_mask=sc.vol_mask, angles=p.start,
                             ctx=ctx)
        sv = np.zeros(3)
        wall = _timed(ctx, lambda: al.grid(grid, sv, mode=3), reps)
        ms, cnt = _ktime(ctx, "xform3d")
        _lib.call("emc_tune_get", ctx.handle, b"align_grid_batch", n, _lib.C.byref(v))
        bytes_ = len(grid) * 5 * 8 * n ** 3
        per = ms / reps
        t0 = time.perf_counter()
        align_particle(al, grid, phi_inc_of=None, theta_inc=3.0, psi_inc=3.0, refine=True)
        ctx.synchronize()
        full = time.perf_counter() - t0
        _line("plain", wall, {"n": n, "orientations": len(grid), "orientation_block": v.value,
                              "orientations_per_s": round(len(grid) / wall, 1),
                              "xform3d_ms_per_call": round(per, 3), "xform3d_launches_per_call": round(cnt / reps, 1),
                              "xform3d_bytes": bytes_, "xform3d_GBps": round(bytes_ / (per * 1e-3) / 1e9, 1),
                              "xform3d_hbm_frac": round(bytes_ / (per * 1e-3) / 8e12, 4),
                              "align_particle_s": round(full, 3)})
        del al


def bench_wedges(ctx, reps):
    from . import synthetic as S
    from .masking import BH_weightMaskMex_batched
    rng = np.random.default_rng(5)
    out = {}
    for n, count in ((64, 1000), (96, 600), (128, 400)):
        tlts = [S.tilt_table(rng.uniform(50.0, 70.0), 3.0, rng.uniform(1.5, 4.0), 2.5) for _ in range(count)]
        recon = np.array([[1024.0, 1024.0, 512.0], [0.0, 0.0, 0.0]])
        xyz = np.stack([rng.uniform(n, 1024 - n, count), rng.uniform(n, 1024 - n, count),
                        rng.uniform(n, 512 - n, count)], 1)
        dst = DeviceArray((count, n, n, n), np.float32, ctx)
        wall = _timed(ctx, lambda: BH_weightMaskMex_batched((n, n, n), 1, tlts, xyz, recon, out=dst), reps)
        sf, c_sf = _ktime(ctx, "sf3d_many")
        ct, c_ct = _ktime(ctx, "ctf2d_many")
        written = count * n ** 3 * 4
        out[n] = {"particles": count, "tilts": int(sum(len(t) for t in tlts)),
                  "ms_per_call": round(wall * 1e3, 3), "sf3d_ms": round(sf / reps, 3),
                  "sf3d_launches": c_sf // reps, "ctf2d_ms": round(ct / reps, 3),
                  "sf3d_written_GBps": round(written / (sf / reps * 1e-3) / 1e9, 1)}
        del dst
    _line("wedges", sum(v["ms_per_call"] for v in out.values()) * 1e-3, {"classes": out})


def bench_copy(ctx, reps, mib=1024):
    """A device-to-device copy of known size: the calibration kernel of the
    HBM byte counters (FETCH_SIZE, WRITE_SIZE)."""
    src = DeviceArray.from_numpy(np.ones(mib << 18, np.float32), ctx=ctx)
    dst = DeviceArray(src.shape, np.float32, ctx)
    wall = _timed(ctx, lambda: _lib.call("emc_memcpy_d2d", ctx.handle, dst.vp, src.vp, C.c_size_t(src.nbytes)),
                  reps)
    _line("copy", wall, {"bytes_read": src.nbytes, "bytes_written": src.nbytes,
                         "GBps": round(2 * src.nbytes / wall / 1e9, 1)})


def bench_gradient(ctx, reps, table="c64", count=256, n=128):
    """The differentiable-alignment workload: one aligner over ``count``
    particles of n^3, each with its own +-60 deg wedge mask and defocus.

    * gradient check: the adjoint gradient at every particle's coarse-grid
      pick against central differences of the device CC (h = 1e-3 rad,
      1e-2 px), as the largest |grad - fd| over the particle's largest |fd|;
    * refinement: coarse 10-degree local grid over 0.6 x the support, pick,
      8 + 8 quasi-Newton iterations (0.6 x, then the full support), and the
      grid as judge (a particle whose refined CC is below its grid pick's
      keeps the grid pose);
    * dense grid: the 1-degree local grid ([15, 1, 15, 1]) scored over the
      full support, pick.
    Each timed from the prepared particles to the poses on the host."""
    from .batched import BatchAligner, collect, finite_difference_check, judge, refine_stage, search_stage
    from .geometry import eulerSearch
    from .masking import BH_weightMaskMex
    from .workload import accuracy, gradient_batch
    t0 = time.perf_counter()
    sc, parts = gradient_batch(count, n)
    gen_s = time.perf_counter() - t0
    wedges = [BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]], ctx=ctx)
              for p in parts]
    al = BatchAligner(n, sc.ref, sc.bandpass, peak_radius=6, table=table, ctx=ctx)
    al.set_particles([p.data for p in parts], wedges, [p.start for p in parts])
    del wedges
    coarse = eulerSearch("C1", 15, 10, 15, 10)
    dense = eulerSearch("C1", 15, 1, 15, 1)

    # the coarse picks (also the judge's poses and the gradient-check poses)
    h0 = search_stage(al, coarse, support_frac=0.6)
    _, sg, _, best = collect(h0)
    g, fd = finite_difference_check(al, coarse[best], sg)
    scale = np.maximum(np.max(np.abs(fd[:, :3]), axis=1), 1e-30)
    scale_s = np.maximum(np.max(np.abs(fd[:, 3:]), axis=1), 1e-30)
    err_a = np.max(np.abs(g[:, :3] - fd[:, :3]), axis=1) / scale
    err_s = np.max(np.abs(g[:, 3:] - fd[:, 3:]), axis=1) / scale_s

    def refined_judged():
        h = search_stage(al, coarse, support_frac=0.6)
        R0 = h["R"].copy()
        s0 = h["s"].copy()
        refine_stage(h, 16, coarse=(0.6, 8))
        R, s, cc, _ = collect(h)
        return judge(al, R0.get(), s0.get(), R, s, cc)

    def dense_grid():
        return collect(search_stage(al, dense, support_frac=1.0))

    res = {}
    for name, fn in (("refine", refined_judged), ("dense", dense_grid)):
        out = fn()                                          # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        ctx.synchronize()
        res[name] = ((time.perf_counter() - t0) / reps, out)
    (tr, (R, s, cc, back)), (td, (Rd, sd, ccd, _)) = res["refine"], res["dense"]
    acc_r = accuracy(parts, R, s)
    acc_d = accuracy(parts, Rd, sd)
    _line("gradient", tr, {
        "n": n, "particles": count, "table": table, "generation_s": round(gen_s, 1),
        "gradient_check": {"h_rad": 1e-3, "h_px": 1e-2,
                           "angle_rel_err_median": round(float(np.median(err_a)), 5),
                           "angle_rel_err_max": round(float(np.max(err_a)), 5),
                           "shift_rel_err_median": round(float(np.median(err_s)), 5),
                           "shift_rel_err_max": round(float(np.max(err_s)), 5)},
        "refinement": {"orientations": len(coarse), "iterations": 16, "ms": round(tr * 1e3, 2),
                       "alignments_per_s": round(count / tr, 1), "judge_fallbacks": int(np.sum(back)),
                       **acc_r},
        "dense_grid": {"orientations": len(dense), "ms": round(td * 1e3, 2),
                       "alignments_per_s": round(count / td, 1), **acc_d},
        "speedup_vs_dense": round(td / tr, 2)})
    del al


def bench_pipeline(ctx, reps, table="c64", tomo=(1024, 1024, 512), edge=64, count=600, snr=0.03, step=7.0,
                   n_peaks=1000, window=96, max_orient=0, plain_sample=8, pixel=10.0, pruned=-1):
    """Template search -> peaks -> windows -> batched refinement (one pass;
    ``reps`` is unused: the search alone runs minutes). ``max_orient`` > 0
    searches only the first that many orientations of the grid (a timing
    sample; recovery is then not meaningful). ``pruned`` 0 / 1 / 2 forces the
    template spectrum's form (padded-volume R2C / pruned DFT / pruned DFT
    with the z stage on the matrix cores); -1 keeps the library's choice."""
    from . import synthetic as S
    from .align_raw import ParticleAligner, align_particle
    from .geometry import eulerSearch
    from .masking import BH_bandpass3d
    from .pipeline import refine_peaks, score_against_truth, search_peaks, synthetic_tomogram, whole_chunk
    from .workload import local_grid, local_matrix, render_fast
    rng = np.random.default_rng(404)
    blobs = S.make_blobs(rng, 30, radius=0.3 * edge, sigma=(2.0, 4.5))
    t0 = time.perf_counter()
    vol, inserted = synthetic_tomogram(405, tomo, blobs, edge, count, snr)
    gen_s = time.perf_counter() - t0
    template = render_fast(blobs, edge).astype(np.float32)
    grid = eulerSearch("C1", 180, step, 180, step)
    n_grid = len(grid)
    if max_orient:
        grid = grid[:max_orient]
    print(f"[bench_paths] pipeline: tomogram {tomo}, {len(inserted)} copies, {len(grid)} of {n_grid} orientations, "
          f"chunk {whole_chunk(tomo, (edge,) * 3)[0]}", file=sys.stderr, flush=True)
    if pruned >= 0:
        _lib.call("emc_tune_set", ctx.handle, b"tsearch_pruned", 0, int(pruned))
    t_start = time.perf_counter()

    def progress(done, total):
        print(f"[bench_paths] pipeline: theta block {done}/{total} searched, "
              f"{time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)

    peaks, ts, tp = search_peaks(vol, template, grid, n_peaks, 22, ctx=ctx, progress=progress)
    print(f"[bench_paths] pipeline: search {ts:.1f} s, {len(peaks)} peaks in {tp:.2f} s", file=sys.stderr, flush=True)
    reference = (-render_fast(blobs, window)).astype(np.float32)   # the copies' contrast (dark)
    bp = BH_bandpass3d((window,) * 3, 1e-3, 400.0, 2.5 * pixel, "GPU", pixel)
    R, centres, cc, valid, back, tr = refine_peaks(vol, peaks, reference, bp, window=window, table=table, ctx=ctx,
                                                   mask_radius=26)
    starts = np.stack([local_matrix(p[3], p[4], p[5] + p[3]) for p in peaks])
    refined = score_against_truth(inserted, R, centres)
    searched = score_against_truth(inserted, starts, peaks[:, :3])
    # the plain per-particle path on a sample of the windows: coarse local
    # grid (1,045 orientations) + the reference's refinement grid
    from .subtomo import ResidentTomogram, extract_subtomos
    res = ResidentTomogram.from_numpy(vol, ctx)
    k = min(plain_sample, len(peaks))
    wins, _, _ = extract_subtomos(res, np.rint(peaks[:k, :3]), (window,) * 3, 0.4 * window, taper=True)
    wh = wins.get()
    del wins, res
    pm, vm = S.sphere_mask(window, window // 8 + 4), S.sphere_mask(window, 0.45 * window)
    bph = bp.get()
    ones = np.ones((window,) * 3, np.float32)                 # no missing wedge in this tomogram
    lg = local_grid()
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(k):
        al = ParticleAligner(wh[i], reference, ones, bph, pm, vol_mask=vm, angles=starts[i])
        align_particle(al, lg, phi_inc_of=None, theta_inc=3.0, psi_inc=3.0, refine=True)
    ctx.synchronize()
    plain_per = (time.perf_counter() - t0) / max(k, 1)
    _line("pipeline", ts + tp + tr, {
        "tomogram": list(tomo), "template": edge, "copies": len(inserted), "snr": snr, "generation_s": round(gen_s, 1),
        "search": {"orientations": len(grid), "grid_orientations": n_grid, "step_deg": step, "pruned": pruned,
                   "chunk": whole_chunk(tomo, (edge,) * 3)[0], "s": round(ts, 2),
                   "ms_per_orientation": round(ts / len(grid) * 1e3, 3)},
        "peaks": {"count": int(len(peaks)), "s": round(tp, 3)},
        "refinement": {"window": window, "particles": int(len(peaks)), "valid_windows": int(valid.sum()),
                       "s": round(tr, 3), "alignments_per_s": round(len(peaks) / tr, 1),
                       "judge_fallbacks": int(back.sum())},
        "against_truth_refined": refined, "against_truth_search_only": searched,
        "plain_grid": {"sample": k, "orientations": len(lg), "s_per_particle": round(plain_per, 3),
                       "alignments_per_s": round(1.0 / plain_per, 3)},
        "refined_vs_plain_throughput": round((len(peaks) / tr) * plain_per, 1)})


PATHS = {"extract": bench_extract, "average": bench_average, "peaks": bench_peaks, "tsprep": bench_tsprep,
         "tsearch": bench_tsearch, "adjoint": bench_adjoint, "refine": bench_refine, "search": bench_search,
         "copy": bench_copy, "wedges": bench_wedges, "plain": bench_plain, "gradient": bench_gradient,
         "pipeline": bench_pipeline}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--path", default="all", choices=["all", *PATHS])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--table", choices=["c64", "f16"], default="c64",
                    help="reference table format of the adjoint and search paths")
    ap.add_argument("--max-orient", type=int, default=0,
                    help="pipeline: search only the first N orientations of the grid (timing sample)")
    ap.add_argument("--step", type=float, default=7.0, help="pipeline: angular step of the search grid (deg)")
    ap.add_argument("--pruned", type=int, default=-1, choices=[-1, 0, 1, 2],
                    help="pipeline: template spectrum by pruned DFT (1; 2 with the z stage on the matrix cores) "
                         "or padded R2C (0); -1 = library default")
    a = ap.parse_args(argv)
    ctx = default_context()
    for name in ([p for p in PATHS if p not in ("gradient", "pipeline")] if a.path == "all" else [a.path]):
        print(f"[bench_paths] {name}", file=sys.stderr, flush=True)
        if name in ("adjoint", "refine", "search", "gradient"):
            PATHS[name](ctx, a.reps, table=a.table)
        elif name == "pipeline":
            PATHS[name](ctx, a.reps, table=a.table, max_orient=a.max_orient, step=a.step, pruned=a.pruned)
        else:
            PATHS[name](ctx, a.reps)


if __name__ == "__main__":
    main()
