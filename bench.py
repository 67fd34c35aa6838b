#!/usr/bin/env python3
"""Benchmark: subtomogram alignments/s on one MI355X.

Workload (one "step" = one batch of particles through alignment): 2,000
subtomograms of mixed box sizes (64^3/96^3/128^3 in 50/30/20 proportion,
FP32 arithmetic, 2.5 A/px), every one distinct: its own object pose (true
pose inside a +-15 deg local search around a random start), shift
(U[-3,3] px), missing wedge (max tilt U[50,70] deg), defocus (U[1.5,4.0]
um), SNR 0.1. The particles of each box size lie in an int16 tomogram
resident in HBM (one particle per cell). A step does, per box size, what the
reference does per particle (alignment/BH_alignRaw3d_v2.m:1036-1094 and the
orientation loops after it): cut every window from the tomogram
(emc_extract_subtomos, 'singleTaper'), build every particle's wedge from its
own tilt table (emc_weight_mask_batched), the reference table, the particle
spectra, the coarse search, the pick and the adjoint refinement. ``value``
counts particles aligned per second; the accuracy block reports how many met
<=1 deg and <=0.5 px against ground truth (``value_within_bar`` counts only
those). The plain GPU path (the drop-in two-stage grid search over the
1,045-orientation local grid, then the reference's refinement) is timed on a
subset of the same particles for the speed-up at equal accuracy.

The timed steps run on one stream with eager refinement launches, so HIP
events time every kernel of the step on its own; the roofline entry is the
kernel with the largest time per step. ``variants`` repeats the step with the
refinement replayed from HIP graphs and with the fp16 reference table.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "subtomo alignments/s at <=1 deg & <=0.5 px error; batched+adjoint vs plain GPU"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
COUNTERS = os.path.join(ROOT, "profiles", "align_eval_grad_counters.md")
TIMED_KERNELS = ("align_eval_grad", "align_fused", "align_final", "prep_particle", "sf3d_many", "ctf2d_many",
                 "extract_subtomos")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=2000, help="particles per step (batched path)")
    ap.add_argument("--plain", type=int, default=10, help="particles timed on the plain path (0 = skip)")
    ap.add_argument("--refine-iters", type=int, default=16)
    ap.add_argument("--coarse-frac", type=float, default=0.6,
                    help="support fraction of the first refinement stage (0 = one full-support stage)")
    ap.add_argument("--coarse-iters", type=int, default=8, help="iterations of the first refinement stage")
    ap.add_argument("--search-frac", type=float, default=0.6,
                    help="support fraction the coarse grid is scored over")
    ap.add_argument("--cpu-sample", type=float, default=20.0,
                    help="seconds budget of the CPU oracle sample (0 = skip)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--table", choices=["c64", "f16"], default="c64",
                    help="reference table format of the batched path")
    ap.add_argument("--no-variant", dest="variant", action="store_false",
                    help="skip the runs with HIP graphs and with the other table format")
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads generating the synthetic particles")
    return ap.parse_args()


# ------------------------------------------------------------ workload

def generate(a):
    """Distinct particles per box size (seeded per particle, so the data do
    not depend on the thread count) and the int16 tomograms holding them."""
    from emclarity_amd.workload import (Degrader, cell_tomogram, make_particle_fast, make_size_class,
                                        mixed_sizes)
    sizes = mixed_sizes(a.batch)
    counts = {n: sizes.count(n) for n in (128, 96, 64) if sizes.count(n)}
    classes, blobs, deg = {}, {}, {}
    for n in counts:
        classes[n], blobs[n] = make_size_class(n)
        deg[n] = Degrader(n)
    particles, cells = {}, {}
    with cf.ThreadPoolExecutor(max_workers=max(1, a.workers)) as ex:
        for n, c in counts.items():
            def one(i, n=n):
                return make_particle_fast(np.random.default_rng([a.seed, n, i]), classes[n], blobs[n], deg[n])
            particles[n] = list(ex.map(one, range(c)))
            cells[n] = cell_tomogram(particles[n], n)
            print(f"[bench] {c} particles of {n}^3 in a {cells[n].data.shape[::-1]} int16 tomogram",
                  file=sys.stderr, flush=True)
    return classes, particles, cells


class BatchedStep:
    """One step of the batched path over every box size on one stream:
    extraction, wedges, reference table, particle spectra, coarse search,
    pick, two-stage adjoint refinement."""

    def __init__(self, ctx, classes, particles, cells, grid, iters, coarse, search_frac, table, graphs):
        from emclarity_amd import _lib
        from emclarity_amd.batched import support_radius_of
        from emclarity_amd.device import DeviceArray
        from emclarity_amd.subtomo import ResidentTomogram
        self.ctx, self.table = ctx, table
        self.grid, self.iters, self.coarse, self.search_frac = grid, iters, coarse, search_frac
        self.order = sorted(particles, reverse=True)      # largest box first
        self.cells = cells
        self.tomo, self.win, self.wed, self.wv, self.dv, self.starts, self.res = {}, {}, {}, {}, {}, {}, {}
        for n in self.order:
            P = len(particles[n])
            self.tomo[n] = ResidentTomogram.from_numpy(cells[n].data, ctx)
            self.win[n] = DeviceArray((P, n, n, n), np.float32, ctx)
            self.wed[n] = DeviceArray((P, n, n, n), np.float32, ctx)
            nb = 4 * n ** 3
            self.wv[n] = [self.win[n].view((n, n, n), np.float32, nb * i) for i in range(P)]
            self.dv[n] = [self.wed[n].view((n, n, n), np.float32, nb * i) for i in range(P)]
            self.starts[n] = np.stack([p.start for p in particles[n]])
            self.res[n] = (DeviceArray.from_numpy(classes[n].ref, ctx=ctx),
                           DeviceArray.from_numpy(classes[n].bandpass, ctx=ctx),
                           support_radius_of(classes[n].bandpass, n))
            _lib.call("emc_tune_set", ctx.handle, b"refine_graph", n, int(bool(graphs)))
        self.support = {n: self.res[n][2] for n in self.order}

    def run(self):
        from emclarity_amd.batched import BatchAligner, collect, refine_stage, search_stage
        from emclarity_amd.masking import BH_weightMaskMex_batched
        from emclarity_amd.subtomo import extract_subtomos
        handles = {}
        for n in self.order:
            c = self.cells[n]
            extract_subtomos(self.tomo[n], c.centres, (n, n, n), 0.4 * n, out=self.win[n])
            BH_weightMaskMex_batched((n, n, n), 1, c.tlt, c.centres, c.recon, out=self.wed[n])
            ref, bp, support = self.res[n]
            al = BatchAligner(n, ref, bp, support_radius=support, peak_radius=6, table=self.table, ctx=self.ctx)
            al.set_particles(self.wv[n], self.dv[n], self.starts[n])
            h = search_stage(al, self.grid, support_frac=self.search_frac)
            handles[n] = refine_stage(h, self.iters, self.coarse)
        out = {}
        for n, h in handles.items():
            R, s, _, _ = collect(h)
            out[n] = (R, s)
        return out

    def eval_launches(self):
        """(support radius, evaluations) of one step's refinement per box size."""
        k = min(self.coarse[1], self.iters) if self.coarse else 0
        stages = []
        if k > 0:
            stages.append((self.coarse[0], k + 1))
        if self.iters - k > 0:
            stages.append((1.0, self.iters - k + 1))
        return {n: [(f * self.support[n], e) for f, e in stages] for n in self.order}


def accuracy(parts, Rs, shifts):
    from emclarity_amd.workload import accuracy as acc
    return acc(parts, Rs, shifts)


def step_accuracy(particles, out):
    parts, Rs, ss = [], [], []
    for n, (R, s) in out.items():
        parts += particles[n]
        Rs += list(R)
        ss += list(s)
    return accuracy(parts, Rs, ss)


def phi_steps(grid):
    steps = {}
    for th in np.unique(grid[:, 1]):
        ph = np.unique(grid[grid[:, 1] == th, 0])
        steps[float(th)] = float(ph[1] - ph[0]) if len(ph) > 1 else 0.0
    return steps


class PlainRun:
    """The drop-in per-particle grid search on the extracted windows and the
    particles' own wedges (device-resident inputs)."""

    def __init__(self, items, classes, grid):
        from emclarity_amd.align_raw import ParticleAligner
        self.grid = grid
        self.psteps = phi_steps(grid)
        self.aligners = [ParticleAligner(win, classes[p.n].ref, wedge, classes[p.n].bandpass, classes[p.n].peak_mask,
                                         vol_mask=classes[p.n].vol_mask, angles=p.start) for p, win, wedge in items]

    def run(self):
        from emclarity_amd.align_raw import align_particle
        return [align_particle(al, self.grid, phi_inc_of=lambda th: self.psteps.get(float(th), 0.0),
                               theta_inc=3.0, psi_inc=3.0, refine=True).row for al in self.aligners]


def plain_accuracy(parts, rows):
    """Plain-path rows [ref, pIdx, phi, theta, psi-phi, CCC, 1, dx, dy, dz]."""
    from emclarity_amd.workload import local_matrix
    Rs = [p.start @ local_matrix(r[2], r[3], r[4] + r[2]) for p, r in zip(parts, rows)]
    return accuracy(parts, Rs, [np.asarray(r[7:10]) for r in rows])


# ------------------------------------------------------------ measurement

def timed(ctx, fn, steps, label):
    """K steps between synchronisations of the stream; HIP-event kernel
    timing on it."""
    from emclarity_amd import _lib
    _lib.call("emc_timing_reset", ctx.handle)
    _lib.call("emc_timing_enable", ctx.handle, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    out = None
    for s in range(steps):
        out = fn()
        print(f"[bench] {label} step {s + 1}/{steps} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    _lib.call("emc_timing_enable", ctx.handle, 0)
    return out, dt


def kernel_times(ctx, steps):
    """Per timed kernel: ms per step, launches per step, average us."""
    import ctypes as C
    from emclarity_amd import _lib
    out = {}
    for name in TIMED_KERNELS:
        ms, cnt = C.c_double(), C.c_int64()
        _lib.call("emc_timing_read", ctx.handle, name.encode(), C.byref(ms), C.byref(cnt))
        if cnt.value:
            out[name] = {"ms_per_step": round(ms.value / steps, 3), "launches_per_step": round(cnt.value / steps, 2),
                         "avg_us": round(ms.value * 1e3 / cnt.value, 2), "_ms": ms.value, "_n": cnt.value}
    return out


def counter_traffic(table):
    """HBM bytes per align_eval_grad launch (FETCH_SIZE + WRITE_SIZE, each
    scaled by the copy calibration) from the committed counter summary, one
    line 'traffic <table> <bytes per launch>' per table format."""
    try:
        with open(COUNTERS) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == table:
                    return int(float(parts[2]))
    except OSError:
        pass
    return None


def eval_roofline(step, kt, steps, table):
    """align_eval_grad_kernel against the HBM roofline: algorithmic bytes of
    each launch (every particle's Q and N at each support frequency, the
    support list and the table region the samples reach, once), summed over
    the timed launches, over their summed HIP-event time."""
    from emclarity_amd.bench_paths import eval_grad_bytes
    k = kt.get("align_eval_grad")
    if not k:
        return None
    total_b, total_s, launches = 0, 0, 0
    for n, stages in step.eval_launches().items():
        P = len(step.wv[n])
        for radius, evals in stages:
            b, s = eval_grad_bytes(n, P, radius, table)
            total_b += b * evals
            total_s += s * evals
            launches += evals
    if launches * steps != k["_n"]:
        raise RuntimeError(f"align_eval_grad launches {k['_n']} != expected {launches * steps}")
    sec = k["_ms"] * 1e-3
    achieved = total_b * steps / sec / 1e9
    traffic = counter_traffic(table)
    gather = (32 if table == "c64" else 16) * 2 * total_s * steps / sec / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": f"align_eval_grad_kernel<{table} table>", "avg_us": k["avg_us"], "launches": k["_n"],
            "bytes_per_launch": round(total_b / launches),
            "gsamples_per_s": round(total_s * steps / sec / 1e9, 2),
            "table_gather_GBps": round(gather, 1),
            "traffic_source": (os.path.relpath(COUNTERS, ROOT) if traffic else None),
            "note": ("algorithmic bytes: Q (8 B) + N (4 B) per particle and support frequency, the support list "
                     "and the reached table region once; each sample also gathers its 2x2x2 table footprint "
                     "(64 B complex64, 32 B fp16) from L2/MALL (table_gather_GBps)")}


def search_kernel(step, kt, steps):
    """The coarse search's pass kernel: FP32 rate (secondary figure)."""
    from emclarity_amd.batched import fused_flops_per_unit, fused_samples_per_unit
    k = kt.get("align_fused")
    if not k:
        return None
    flops = samples = 0
    for n in step.order:
        u = len(step.wv[n]) * len(step.grid)
        r = step.search_frac * step.support[n]
        flops += fused_flops_per_unit(n, r, 6) * u
        samples += fused_samples_per_unit(n, r) * u
    sec = k["_ms"] * 1e-3
    return {"kernel": "align_pass_kernel<6>", "avg_us": k["avg_us"], "launches": k["_n"],
            "achieved_tflops": round(flops * steps / sec / 1e12, 3),
            "frac_fp32_peak": round(flops * steps / sec / 1e12 / FP32_PEAK_TFLOPS, 4),
            "gsamples_per_s": round(samples * steps / sec / 1e9, 2)}


def cpu_baseline(budget_s, classes, items):
    """The oracle's FP32 C restatement of the reference's two-stage plain
    alignment body (oracle/emc_cpu.c, OpenMP over orientations, all the
    threads OpenMP is given) on a bounded sample: batches of orientations per
    size class within the time budget, extrapolated to the 1,045 + 45 + 1
    orientations (local grid, refinement grid, final) a particle needs.
    items: {n: (particle, window (host), wedge (host))}."""
    from oracle import emc_cpu
    threads = emc_cpu.max_threads()
    per_orient, counts = {}, {}
    t_all = time.perf_counter()
    for n in (64, 96, 128):
        if n not in items:
            continue
        p, win, wedge = items[n]
        sc = classes[n]
        share = budget_s * {64: 0.2, 96: 0.3, 128: 0.5}[n]
        k, elapsed = 0, 0.0
        while elapsed < share:
            pts = np.array([[3.0 * (k + j), 3.0, -3.0 * (k + j)] for j in range(2 * threads)])
            t0 = time.perf_counter()
            emc_cpu.align_grid(win, sc.ref, sc.vol_mask, sc.peak_mask, sc.bandpass, wedge, p.start, pts)
            elapsed += time.perf_counter() - t0
            k += len(pts)
        per_orient[n] = elapsed / k
        counts[n] = k
    n_orient = 1045 + 45 + 1
    mix = {64: 0.5, 96: 0.3, 128: 0.2}
    sec = sum(mix[n] * per_orient[n] * n_orient for n in per_orient) / sum(mix[n] for n in per_orient)
    return {"value": round(1.0 / sec, 6), "unit": "alignments/s", "cores": threads, "kind": "port",
            "sample": (f"FP32 C restatement (oracle/emc_cpu.c) of the reference's two-stage alignRaw body, "
                       f"OpenMP over orientations on {threads} threads: "
                       + "/".join(str(counts.get(n, 0)) for n in (64, 96, 128))
                       + f" orientations at 64/96/128^3 ({time.perf_counter() - t_all:.1f} s), extrapolated to "
                       f"{n_orient} orientations per particle, 50/30/20 mix")}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    from emclarity_amd.device import default_context
    from emclarity_amd.geometry import eulerSearch
    from emclarity_amd.masking import BH_weightMaskMex
    from emclarity_amd.workload import local_grid

    ctx = default_context()
    t_gen = time.perf_counter()
    classes, particles, cells = generate(a)
    print(f"[bench] generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    total = sum(len(v) for v in particles.values())
    # the 10-degree coarse grid of the local search ([theta_max 15, step 10,
    # psi_max 15, step 10]: 28 orientations); the adjoint refinement takes
    # the best of them to the final pose, in two stages (the first
    # iterations over a smaller support, the rest over the full support)
    coarse = eulerSearch("C1", 15, 10, 15, 10)
    c2f = (a.coarse_frac, a.coarse_iters) if a.coarse_frac > 0 and a.coarse_iters > 0 else None

    def batched(table, graphs, label):
        step = BatchedStep(ctx, classes, particles, cells, coarse, a.refine_iters, c2f, a.search_frac, table, graphs)
        for _ in range(a.warmup):
            step.run()
        out, dt = timed(ctx, step.run, a.steps, label)
        kt = kernel_times(ctx, a.steps)
        return step, {"value": total * a.steps / dt, "dt": dt, "acc": step_accuracy(particles, out), "kt": kt}

    step, main_run = batched(a.table, False, f"batched ({a.table} table)")
    kt = main_run["kt"]
    roof = eval_roofline(step, kt, a.steps, a.table)
    search = search_kernel(step, kt, a.steps)
    dominant = max(kt, key=lambda k: kt[k]["_ms"]) if kt else None
    variants = None
    if a.variant:
        variants = {}
        _, g = batched(a.table, True, f"batched ({a.table} table, HIP graphs)")
        variants["hip_graph_refinement"] = {"value": round(g["value"], 3), "ms_per_step": round(g["dt"] * 1e3 / a.steps, 2),
                                            "accuracy": g["acc"]}
        other = "f16" if a.table == "c64" else "c64"
        ostep, o = batched(other, False, f"batched ({other} table)")
        variants[f"{other}_table"] = {"value": round(o["value"], 3), "ms_per_step": round(o["dt"] * 1e3 / a.steps, 2),
                                      "accuracy": o["acc"], "roofline": eval_roofline(ostep, o["kt"], a.steps, other)}
        del ostep

    # a few particles of each size: their extracted windows and own wedges
    pick = {64: 5, 96: 3, 128: 2}
    items = []
    for n in step.order:
        c = cells[n]
        for i in range(min(pick.get(n, 0) * max(a.plain, 0) // 10, len(particles[n]))):
            w = BH_weightMaskMex((n, n, n), 1, c.tlt[i], c.centres[i], c.recon, ctx=ctx)
            items.append((particles[n][i], step.wv[n][i], w))
    plain = None
    if a.plain > 0 and items:
        pr = PlainRun(items, classes, local_grid())
        pr.run()
        rows, pdt = timed(ctx, pr.run, 1, "plain")
        plain = {"value": round(len(items) / pdt, 4), "unit": "alignments/s", "particles": len(items),
                 "orientations": 1045, "accuracy": plain_accuracy([p for p, _, _ in items], rows)}
    cpu = None
    if a.cpu_sample > 0 and items:
        first = {}
        for p, win, w in items:
            first.setdefault(p.n, (p, win.get(), w.get()))
        cpu = cpu_baseline(a.cpu_sample, classes, first)

    value, dt, acc = main_run["value"], main_run["dt"], main_run["acc"]
    kernels = {k: {kk: vv for kk, vv in v.items() if not kk.startswith("_")} for k, v in kt.items()}
    line = {"metric": METRIC, "value": round(value, 3), "unit": "alignments/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt * 1e3 / a.steps, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "batched alignment: 2,000 distinct subtomograms of mixed box size 64/96/128^3 "
                                   "(50/30/20), each its own pose, shift, wedge U[50,70] deg and defocus "
                                   "U[1.5,4.0] um, SNR 0.1, +-15 deg local search; windows cut from resident "
                                   "int16 tomograms and wedges built inside the step",
                       "particles_per_step": total, "path": "batched+adjoint", "reference_table": a.table,
                       "streams": 1, "refinement_launch": "eager (HIP events per kernel)",
                       "coarse_orientations": int(len(coarse)), "search_support_fraction": a.search_frac,
                       "refine_iterations": a.refine_iters,
                       "refine_schedule": (f"{c2f[1]} at {c2f[0]:g} x the support radius, then {a.refine_iters - c2f[1]} "
                                           f"over the full support") if c2f else "full support"},
            "accuracy": acc, "value_within_bar": round(value * acc["fraction"], 3),
            "kernels_per_step": kernels, "dominant_kernel": dominant,
            "variants": variants, "plain_gpu": plain,
            "speedup_vs_plain": round(value / plain["value"], 2) if plain else None,
            "roofline": roof, "search_kernel": search, "cpu_baseline": cpu}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
