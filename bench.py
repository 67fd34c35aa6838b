#!/usr/bin/env python3
"""Benchmark: subtomogram alignments/s on one MI355X.

Workload (one "step" = one batch of particles through the alignment):
mixed-size subtomograms (64^3/96^3/128^3 in 50/30/20 proportion, FP32,
2.5 A/px), each with its own missing wedge (max tilt U[50,70] deg) and
defocus (U[1.5,4.0] um), SNR 0.1, true pose inside a +-15 deg local search of
1,045 orientations around a perturbed start. Inputs are resident in HBM
before the timed region. ``value`` counts particles aligned per second; the
accuracy block reports how many met <=1 deg and <=0.5 px against ground truth.

Paths: ``plain`` is the drop-in per-particle grid search (the two-stage
alignRaw body for every orientation, then the reference's local refinement
and final translational pass).

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
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=10, help="particles per step")
    ap.add_argument("--path", default="plain", choices=["plain"])
    ap.add_argument("--cpu-sample", type=float, default=20.0,
                    help="seconds budget of the CPU oracle sample (0 = skip)")
    ap.add_argument("--seed", type=int, default=2024)
    return ap.parse_args()


def phi_steps(grid):
    steps = {}
    for th in np.unique(grid[:, 1]):
        ph = np.unique(grid[grid[:, 1] == th, 0])
        steps[float(th)] = float(ph[1] - ph[0]) if len(ph) > 1 else 0.0
    return steps


class PlainBatch:
    """Per-particle plain GPU grid search over a batch (device-resident)."""

    def __init__(self, particles, classes, grid):
        from emclarity_amd.align_raw import ParticleAligner
        from emclarity_amd.masking import BH_weightMaskMex
        self.grid = grid
        self.psteps = phi_steps(grid)
        self.aligners = []
        for p in particles:
            sc = classes[p.n]
            n = p.n
            wedge = BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]])
            self.aligners.append(ParticleAligner(p.data, sc.ref, wedge, sc.bandpass, sc.peak_mask,
                                                 vol_mask=sc.vol_mask, angles=p.start))

    def run(self):
        from emclarity_amd.align_raw import align_particle
        out = []
        for al in self.aligners:
            r = align_particle(al, self.grid, phi_inc_of=lambda th: self.psteps.get(float(th), 0.0),
                               theta_inc=3.0, psi_inc=3.0, refine=True)
            out.append(r.row)
        return out


def accuracy(particles, rows):
    from emclarity_amd.workload import angle_error_deg, local_matrix
    ok = 0
    errs = []
    for p, row in zip(particles, rows):
        R = p.start @ local_matrix(row[2], row[3], row[4] + row[2])
        ea = angle_error_deg(R, p.truth_R)
        es = float(np.max(np.abs(np.asarray(row[7:10]) - p.shift)))
        errs.append((ea, es))
        ok += int(ea <= 1.0 and es <= 0.5)
    e = np.array(errs)
    return {"within_1deg_0p5px": ok, "of": len(rows), "median_angle_err_deg": float(np.median(e[:, 0])),
            "median_shift_err_px": float(np.median(e[:, 1]))}


def cpu_baseline(budget_s, classes, particles):
    """FP64 oracle (CPU restatement of the two-stage body) on a bounded
    sample: a few orientations per size class, extrapolated to the 1,045 +
    refinement orientations each particle needs."""
    from oracle import emc_oracle as O
    from emclarity_amd.masking import BH_weightMaskMex
    per_orient = {}
    t_all = time.perf_counter()
    counts = {}
    for n in (64, 96, 128):
        p = next(q for q in particles if q.n == n)
        sc = classes[n]
        wedge = BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]]).get()
        inp = O.AlignInputs(particle=p.data.astype(np.float64), ref=sc.ref.astype(np.float64),
                            ref_wgt_rot=None, ref_wgt=1.0,
                            ref_ft1=O.ref_ft1_of(sc.ref.astype(np.float64), sc.vol_mask, [[0] * 3, [0] * 3]),
                            vol_mask=sc.vol_mask.astype(np.float64), peak_mask=sc.peak_mask.astype(np.float64),
                            bandpass=sc.bandpass.astype(np.float64), wedge=wedge.astype(np.float64),
                            angles=p.start, shift_val=np.zeros(3), pad_calc=[[0] * 3, [0] * 3])
        k = 0
        t0 = time.perf_counter()
        share = budget_s * {64: 0.2, 96: 0.3, 128: 0.5}[n]
        while True:
            R = p.start @ O.define_matrix([3.0 * k, 3.0, 0.0 - 3.0 * k], "Bah", "inv")
            O.align_orientation(inp, R, np.zeros(3))
            k += 1
            if time.perf_counter() - t0 > share:
                break
        per_orient[n] = (time.perf_counter() - t0) / k
        counts[n] = k
    n_orient = 1045 + 45 + 1
    mix = {64: 0.5, 96: 0.3, 128: 0.2}
    sec_per_particle = sum(mix[n] * per_orient[n] * n_orient for n in mix)
    return {"value": 1.0 / sec_per_particle, "unit": "alignments/s", "cores": 1, "kind": "port",
            "sample": (f"FP64 numpy oracle, two-stage alignRaw body: {counts[64]}/{counts[96]}/{counts[128]} "
                       f"orientations at 64/96/128^3 ({time.perf_counter() - t_all:.1f} s), extrapolated to "
                       f"{n_orient} orientations per particle, 50/30/20 mix; single thread")}


def kernel_time(ctx, name):
    import ctypes as C
    from emclarity_amd import _lib
    ms = C.c_double()
    cnt = C.c_int64()
    _lib.call("emc_timing_read", ctx.handle, name.encode(), C.byref(ms), C.byref(cnt))
    return ms.value, cnt.value


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {a.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    from emclarity_amd import _lib
    from emclarity_amd.device import default_context
    from emclarity_amd.workload import local_grid, make_particle, make_size_class, mixed_sizes

    ctx = default_context()
    rng = np.random.default_rng(a.seed)
    classes, blobs = {}, {}
    for n in (64, 96, 128):
        classes[n], blobs[n] = make_size_class(n)
    sizes = mixed_sizes(a.batch)
    particles = [make_particle(rng, classes[n], blobs[n]) for n in sizes]
    grid = local_grid()
    print(f"[bench] {len(particles)} particles {dict(zip(*np.unique(sizes, return_counts=True)))}, "
          f"{len(grid)} orientations, path={a.path}", file=sys.stderr, flush=True)
    runner = PlainBatch(particles, classes, grid)
    for w in range(a.warmup):
        rows = runner.run()
        print(f"[bench] warmup {w + 1}/{a.warmup} done", file=sys.stderr, flush=True)
    _lib.call("emc_timing_reset", ctx.handle)
    _lib.call("emc_timing_enable", ctx.handle, 1)
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in range(a.steps):
        rows = runner.run()
        print(f"[bench] step {s + 1}/{a.steps} done", file=sys.stderr, flush=True)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    _lib.call("emc_timing_enable", ctx.handle, 0)
    acc = accuracy(particles, rows)
    value = len(particles) * a.steps / dt

    # roofline of the rotate kernel (the plain path's most frequent kernel):
    # algorithmic bytes per launch = 4 N_in + 4 N_out (+ 4 N for the volMask
    # multiplier of the stage-2 particle rotation, averaged in by count)
    ms, cnt = kernel_time(ctx, "xform3d")
    vox = sum(p.n ** 3 for p in particles) / len(particles)
    bytes_per_launch = 8.0 * vox * (1.0 + 0.25 / 5)   # 1 of 5 launches reads volMask too
    achieved = bytes_per_launch * cnt / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "xform3d", "avg_us": round(ms * 1e3 / max(cnt, 1), 3), "launches": cnt}
    cpu = cpu_baseline(a.cpu_sample, classes, particles) if a.cpu_sample > 0 else None
    line = {"metric": METRIC, "value": round(value, 4), "unit": "alignments/s", "n_gpus": 1,
            "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(dt * 1e3 / a.steps, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "batched alignment, mixed 64/96/128^3 (50/30/20), per-particle wedge "
                                   "U[50,70] deg + defocus U[1.5,4.0] um, SNR 0.1, 1,045-orientation local "
                                   "search + refinement", "particles_per_step": len(particles),
                       "path": a.path, "orientations": int(len(grid))},
            "accuracy": acc, "roofline": roof, "cpu_baseline": cpu}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
