#!/usr/bin/env python3
"""Benchmark: subtomogram alignments/s on one MI355X.

Workload (one "step" = one batch of particles through alignment): mixed-size
subtomograms (64^3/96^3/128^3 in 50/30/20 proportion, FP32, 2.5 A/px), each
with its own missing wedge (max tilt U[50,70] deg) and defocus (U[1.5,4.0]
um), SNR 0.1, true pose inside a +-15 deg local search around a perturbed
start. Particles and wedges are resident in HBM before the timed region.
``value`` counts particles aligned per second by the batched + adjoint path;
the accuracy block reports how many met <=1 deg and <=0.5 px against ground
truth. The plain GPU path (the drop-in two-stage grid search over the
1,045-orientation local grid, then the reference's refinement) is timed on a
subset of the same particles for the speed-up at equal accuracy.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "subtomo alignments/s at <=1 deg & <=0.5 px error; batched+adjoint vs plain GPU"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=2000, help="particles per step (batched path)")
    ap.add_argument("--pool", type=int, default=100,
                    help="distinct rendered particles; the rest of the batch are copies of them with fresh "
                         "noise and new start orientations")
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
                    help="skip the run with the other reference table format")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="one untimed device-to-device copy of this size first: a known-size "
                         "kernel that calibrates HBM byte counters (FETCH_SIZE, WRITE_SIZE)")
    return ap.parse_args()


def accuracy_rows(particles, rows):
    """Plain-path rows [ref, pIdx, phi, theta, psi-phi, CCC, 1, dx, dy, dz]."""
    from emclarity_amd.workload import local_matrix
    Rs = [p.start @ local_matrix(r[2], r[3], r[4] + r[2]) for p, r in zip(particles, rows)]
    return accuracy(particles, Rs, [np.asarray(r[7:10]) for r in rows])


def accuracy(particles, Rs, shifts):
    from emclarity_amd.workload import angle_error_deg
    errs = np.array([(angle_error_deg(R, p.truth_R), float(np.max(np.abs(s - p.shift))))
                     for p, R, s in zip(particles, Rs, shifts)])
    ok = int(np.sum((errs[:, 0] <= 1.0) & (errs[:, 1] <= 0.5)))
    return {"within_1deg_0p5px": ok, "of": len(particles), "fraction": round(ok / len(particles), 4),
            "median_angle_err_deg": round(float(np.median(errs[:, 0])), 4),
            "median_shift_err_px": round(float(np.median(errs[:, 1])), 4)}


def phi_steps(grid):
    steps = {}
    for th in np.unique(grid[:, 1]):
        ph = np.unique(grid[grid[:, 1] == th, 0])
        steps[float(th)] = float(ph[1] - ph[0]) if len(ph) > 1 else 0.0
    return steps


class PlainRun:
    """The drop-in per-particle grid search (device-resident inputs)."""

    def __init__(self, particles, classes, wedges, grid):
        from emclarity_amd.align_raw import ParticleAligner
        self.grid = grid
        self.psteps = phi_steps(grid)
        self.aligners = [ParticleAligner(p.data, classes[p.n].ref, w, classes[p.n].bandpass,
                                         classes[p.n].peak_mask, vol_mask=classes[p.n].vol_mask,
                                         angles=p.start) for p, w in zip(particles, wedges)]

    def run(self):
        from emclarity_amd.align_raw import align_particle
        return [align_particle(al, self.grid, phi_inc_of=lambda th: self.psteps.get(float(th), 0.0),
                               theta_inc=3.0, psi_inc=3.0, refine=True).row for al in self.aligners]


class BatchedRun:
    """Batched Fourier-space search over a coarse local grid + adjoint
    refinement, one BatchAligner per box size (particles, wedges resident)."""

    def __init__(self, particles, classes, wedges, grid, refine_iters, table="c64", coarse=None, search_frac=1.0):
        self.table = table
        # largest box first: its refinement, the longest, then overlaps the
        # smaller sizes' searches, and the smallest size's short refinement
        # is the step's tail
        self.groups = {}
        for i, p in sorted(enumerate(particles), key=lambda ip: -ip[1].n):
            self.groups.setdefault(p.n, []).append(i)
        for n in self.groups:
            self.groups[n].sort()
        self.particles, self.classes, self.wedges = particles, classes, wedges
        self.grid, self.iters, self.coarse, self.search_frac = grid, refine_iters, coarse, search_frac
        # per size class: reference and bandpass resident in HBM, and the
        # bandpass's support radius (a property of the filter, made once)
        from emclarity_amd.batched import support_radius_of
        from emclarity_amd.device import Context, DeviceArray
        self.resident = {n: (DeviceArray.from_numpy(classes[n].ref), DeviceArray.from_numpy(classes[n].bandpass),
                             support_radius_of(classes[n].bandpass, n)) for n in self.groups}
        # one context (stream) per box size: the size classes' chains of
        # table build, preparation, search, pick and refinement overlap
        self.ctxs = {n: Context(0) for n in self.groups}

    def run(self):
        from emclarity_amd.batched import BatchAligner, collect, refine_stage, search_stage
        R = [None] * len(self.particles)
        S = [None] * len(self.particles)
        self.units = {}
        aligners, handles = {}, {}
        for n, idx in self.groups.items():      # tables and preparation, overlapped
            ref, bp, support = self.resident[n]
            al = BatchAligner(n, ref, bp, support_radius=support, peak_radius=6, table=self.table, ctx=self.ctxs[n])
            al.set_particles([self.particles[i].data for i in idx], [self.wedges[i] for i in idx],
                             [self.particles[i].start for i in idx])
            aligners[n] = al
            # units and the support radius the search scores them over (the
            # flop and sample counts of the roofline follow that radius)
            self.units[n] = (len(idx) * len(self.grid), al.support * self.search_frac)
        prev = None
        for n in self.groups:                   # searches one after another (each fills the GPU)
            handles[n] = search_stage(aligners[n], self.grid, after=prev, support_frac=self.search_frac)
            prev = self.ctxs[n]
        for n in self.groups:                   # refinements overlap the later searches and each other
            refine_stage(handles[n], self.iters, self.coarse)
        for n, idx in self.groups.items():
            r, s, _, _ = collect(handles[n])
            for k, i in enumerate(idx):
                R[i], S[i] = r[k], s[k]
        return R, S


def cpu_baseline(budget_s, classes, particles, host, wedges):
    """The oracle's FP32 C restatement of the reference's two-stage plain
    alignment body (oracle/emc_cpu.c, OpenMP over orientations, all the
    threads OpenMP is given) on a bounded sample: batches of orientations per
    size class within the time budget, extrapolated to the 1,045 + 45 + 1
    orientations (local grid, refinement grid, final) a particle needs."""
    from oracle import emc_cpu
    threads = emc_cpu.max_threads()
    per_orient, counts = {}, {}
    t_all = time.perf_counter()
    for n in (64, 96, 128):
        i = next(k for k, q in enumerate(particles) if q.n == n)
        p, sc = particles[i], classes[n]
        share = budget_s * {64: 0.2, 96: 0.3, 128: 0.5}[n]
        k, elapsed = 0, 0.0
        while elapsed < share:
            pts = np.array([[3.0 * (k + j), 3.0, -3.0 * (k + j)] for j in range(2 * threads)])
            t0 = time.perf_counter()
            emc_cpu.align_grid(host[i], sc.ref, sc.vol_mask, sc.peak_mask, sc.bandpass, wedges[i].get(), p.start,
                               pts)
            elapsed += time.perf_counter() - t0
            k += len(pts)
        per_orient[n] = elapsed / k
        counts[n] = k
    n_orient = 1045 + 45 + 1
    mix = {64: 0.5, 96: 0.3, 128: 0.2}
    sec = sum(mix[n] * per_orient[n] * n_orient for n in mix)
    return {"value": round(1.0 / sec, 6), "unit": "alignments/s", "cores": threads, "kind": "port",
            "sample": (f"FP32 C restatement (oracle/emc_cpu.c) of the reference's two-stage alignRaw body, "
                       f"OpenMP over orientations on {threads} threads: {counts[64]}/{counts[96]}/{counts[128]} "
                       f"orientations at 64/96/128^3 ({time.perf_counter() - t_all:.1f} s), extrapolated to "
                       f"{n_orient} orientations per particle, 50/30/20 mix")}


COUNTERS = os.path.join(ROOT, "profiles", "align_fused_counters.md")


def counter_traffic(table):
    """HBM bytes per align_fused launch (FETCH_SIZE + WRITE_SIZE, calibrated
    against a copy of known size) from the committed counter summary, which
    carries one line 'traffic <table> <bytes per launch>' per table format."""
    try:
        with open(COUNTERS) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == table:
                    return {"bytes_per_launch": int(float(parts[2])),
                            "source": os.path.relpath(COUNTERS, ROOT)}
    except OSError:
        pass
    return None


def kernel_time(ctxs, name):
    """Total ms and launches of a timed kernel over the contexts' streams."""
    import ctypes as C
    from emclarity_amd import _lib
    total, n = 0.0, 0
    for ctx in ctxs:
        ms, cnt = C.c_double(), C.c_int64()
        _lib.call("emc_timing_read", ctx.handle, name.encode(), C.byref(ms), C.byref(cnt))
        total += ms.value
        n += cnt.value
    return total, n


def timed(ctxs, fn, steps, label):
    """K steps between synchronisations of every stream involved; HIP-event
    kernel timing on each of them."""
    from emclarity_amd import _lib
    for ctx in ctxs:
        _lib.call("emc_timing_reset", ctx.handle)
        _lib.call("emc_timing_enable", ctx.handle, 1)
        ctx.synchronize()
    t0 = time.perf_counter()
    out = None
    for s in range(steps):
        out = fn()
        print(f"[bench] {label} step {s + 1}/{steps} done", file=sys.stderr, flush=True)
    for ctx in ctxs:
        ctx.synchronize()
    dt = time.perf_counter() - t0
    for ctx in ctxs:
        _lib.call("emc_timing_enable", ctx.handle, 0)
    return out, dt


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    from emclarity_amd.batched import fused_flops_per_unit, fused_samples_per_unit
    from emclarity_amd.device import DeviceArray, default_context
    from emclarity_amd.geometry import eulerSearch
    from emclarity_amd.masking import BH_weightMaskMex
    from emclarity_amd.workload import copy_particle, local_grid, make_particle, make_size_class, mixed_sizes

    ctx = default_context()
    if a.copy_mib > 0:
        src = DeviceArray.from_numpy(np.ones(a.copy_mib << 18, np.float32))
        dst = src.copy()
        ctx.synchronize()
        del src, dst
    rng = np.random.default_rng(a.seed)
    classes, blobs = {}, {}
    for n in (64, 96, 128):
        classes[n], blobs[n] = make_size_class(n)
    sizes = mixed_sizes(a.batch)
    pool_sizes = mixed_sizes(min(a.pool, a.batch))
    particles = [make_particle(rng, classes[n], blobs[n], keep_signal=True) for n in pool_sizes]
    wedges = [BH_weightMaskMex((p.n,) * 3, 1, p.tlt, [p.n // 2 + 1] * 3, [[p.n] * 3, [0, 0, 0]])
              for p in particles]
    host = [p.data for p in particles]
    for p in particles:
        p.data = DeviceArray.from_numpy(p.data)   # resident before timing
    # the rest of the batch: each pool particle seen again with a fresh noise
    # realisation and a new start (same object, pose, shift, wedge, defocus)
    for n in (64, 96, 128):
        src = [i for i, p in enumerate(particles[:len(pool_sizes)]) if p.n == n]
        for k in range(sizes.count(n) - len(src)):
            i = src[k % len(src)]
            q = copy_particle(rng, particles[i])
            q.data = DeviceArray.from_numpy(q.data)
            particles.append(q)
            wedges.append(wedges[i])
            host.append(None)
    for p in particles:
        p.signal = None
    # the 10-degree coarse grid of the local search ([theta_max 15, step 10,
    # psi_max 15, step 10]: 28 orientations); the adjoint refinement takes
    # the best of them to the final pose
    coarse = eulerSearch("C1", 15, 10, 15, 10)
    # refinement in two stages: the first iterations over a smaller support
    # (fewer frequencies, a smoother score), the rest over the full support
    c2f = (a.coarse_frac, a.coarse_iters) if a.coarse_frac > 0 and a.coarse_iters > 0 else None
    print(f"[bench] {len(particles)} particles {dict(zip(*[x.tolist() for x in np.unique(sizes, return_counts=True)]))} "
          f"({len(pool_sizes)} rendered, the rest fresh-noise copies with new starts), "
          f"coarse grid {len(coarse)} orientations + {a.refine_iters} adjoint iterations",
          file=sys.stderr, flush=True)

    def batched(table):
        run = BatchedRun(particles, classes, wedges, coarse, a.refine_iters, table, c2f, a.search_frac)
        for w in range(a.warmup):
            run.run()
        ctxs = [ctx] + list(run.ctxs.values())
        (Rb, Sb), dt = timed(ctxs, run.run, a.steps, f"batched ({table} table)")
        ms, cnt = kernel_time(ctxs, "align_fused")
        flops = sum(fused_flops_per_unit(n, sup, 6) * u for n, (u, sup) in run.units.items())
        samples = sum(fused_samples_per_unit(n, sup) * u for n, (u, sup) in run.units.items())
        return {"value": len(particles) * a.steps / dt, "dt": dt, "acc": accuracy(particles, Rb, Sb),
                "ms": ms, "launches": cnt, "flops": flops * a.steps, "samples": samples * a.steps,
                "units": run.units}

    main_run = batched(a.table)
    other = "f16" if a.table == "c64" else "c64"
    alt = batched(other) if a.variant else None
    value, dt, acc = main_run["value"], main_run["dt"], main_run["acc"]

    # roofline of the fused search kernel against the FP32 peak; its limit in
    # practice is the trilinear gather traffic past L2, measured by the HBM
    # byte counters (copy-calibrated) in the committed counter summary
    ms, cnt = main_run["ms"], main_run["launches"]
    avg_us = ms * 1e3 / max(cnt, 1)
    achieved = main_run["flops"] / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    traffic = counter_traffic(a.table)
    # the same units through the real-space chain would move
    # B_unit = 4 W^3 + 16 C^3 + 40 H bytes each (H = C^2 (C/2+1))
    units = main_run["units"]
    b_unit = {n: 4 * n ** 3 + 16 * n ** 3 + 40 * n * n * (n // 2 + 1) for n in units}
    eq_gbs = sum(b_unit[n] * u for n, (u, _) in units.items()) * a.steps / (ms * 1e-3) / 1e9
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "kernel": f"align_fused<6, {a.table} table>", "avg_us": round(avg_us, 2), "launches": cnt,
            "flops_per_launch": round(main_run["flops"] / max(cnt, 1)),
            "gsamples_per_s": round(main_run["samples"] / (ms * 1e-3) / 1e9, 2),
            "traffic_GBps": round(traffic["bytes_per_launch"] / (avg_us * 1e-6) / 1e9, 1) if traffic else None,
            "traffic_source": traffic["source"] if traffic else None,
            "note": ("FP32 arithmetic vs the 157.3 TFLOP/s FP32 peak (packed FP32 is disabled in this build, "
                     "so 78.6 is attainable); the kernel is limited by trilinear gathers that miss L2"),
            "realspace_chain_equivalent_GBps": round(eq_gbs, 1)}
    variants = None
    if alt:
        variants = {f"{other}_table": {
            "value": round(alt["value"], 3), "ms_per_step": round(alt["dt"] * 1e3 / a.steps, 2),
            "accuracy": alt["acc"], "align_fused_avg_us": round(alt["ms"] * 1e3 / max(alt["launches"], 1), 2),
            "achieved_tflops": round(alt["flops"] / (alt["ms"] * 1e-3) / 1e12, 3) if alt["ms"] > 0 else None,
            "traffic": (counter_traffic(other) or {}).get("bytes_per_launch")}}

    plain = None
    if a.plain > 0:
        sub = []
        for n, k in ((64, 5), (96, 3), (128, 2)):
            sub += [i for i, p in enumerate(particles) if p.n == n][: max(1, round(a.plain * k / 10))]
        pp = [particles[i] for i in sub]
        pr = PlainRun(pp, classes, [wedges[i] for i in sub], local_grid())
        pr.run()
        rows, pdt = timed([ctx], pr.run, 1, "plain")
        plain = {"value": round(len(pp) / pdt, 4), "unit": "alignments/s", "particles": len(pp),
                 "orientations": 1045, "accuracy": accuracy_rows(pp, rows)}

    cpu = cpu_baseline(a.cpu_sample, classes, particles, host, wedges) if a.cpu_sample > 0 else None
    line = {"metric": METRIC, "value": round(value, 3), "unit": "alignments/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt * 1e3 / a.steps, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "batched alignment: mixed 64/96/128^3 subtomograms (50/30/20), own wedge "
                                   "U[50,70] deg + defocus U[1.5,4.0] um, SNR 0.1, +-15 deg local search",
                       "particles_per_step": len(particles), "distinct_rendered": len(pool_sizes),
                       "path": "batched+adjoint",
                       "reference_table": a.table,
                       "coarse_orientations": int(len(coarse)), "search_support_fraction": a.search_frac,
                       "refine_iterations": a.refine_iters,
                       "refine_schedule": (f"{c2f[1]} at {c2f[0]:g} x the support radius, then {a.refine_iters - c2f[1]} "
                                           f"over the full support") if c2f else "full support"},
            "accuracy": acc, "variants": variants, "plain_gpu": plain,
            "speedup_vs_plain": round(value / plain["value"], 2) if plain else None,
            "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
