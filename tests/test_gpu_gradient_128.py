"""The differentiable-alignment workload: 256 subtomograms of 128^3 at SNR
0.1, each with a +-60 deg wedge and its own defocus U[1.5, 4.0] um, in one
batch.

* The adjoint gradient with respect to the three local Euler angles and the
  three shifts, at every particle's 10-degree coarse-grid pick, against
  central differences of the device score (h = 1e-3 rad, 1e-2 px) for all
  256 particles, and against the FP64 oracle (full frequency grid) for a
  subset of them.
* Gradient refinement from the coarse pick (8 + 8 quasi-Newton iterations,
  the grid as fallback and judge) against the dense 1-degree local grid
  ([15, 1, 15, 1], 23,281 orientations, full support) on 64 of the particles:
  the refinement is at least as accurate against the ground truth, every
  particle's refined CC is at least its grid pick's, and both reach the
  <= 1 deg / <= 0.5 px bar on nearly all particles."""
import numpy as np
import pytest

from emclarity_amd.batched import (BatchAligner, collect, euler_derivatives, euler_eval, finite_difference_check,
                                   judge, refine_stage, search_stage)
from emclarity_amd.geometry import eulerSearch
from emclarity_amd.masking import BH_weightMaskMex
from emclarity_amd.workload import accuracy, gradient_batch
from oracle import emc_oracle as O

pytestmark = pytest.mark.gpu

N = 128
COUNT = 256
ORACLE_SUBSET = [0, 97, 255]
DENSE_SUBSET = 64


@pytest.fixture(scope="module")
def batch(ctx):
    sc, parts = gradient_batch(COUNT, N)
    assert all(p.max_tilt == 60.0 for p in parts)
    wedges = [BH_weightMaskMex((N, N, N), 1, p.tlt, [N // 2 + 1] * 3, [[N, N, N], [0, 0, 0]], ctx=ctx)
              for p in parts]
    al = BatchAligner(N, sc.ref, sc.bandpass, peak_radius=6, ctx=ctx)
    al.set_particles([p.data for p in parts], wedges, [p.start for p in parts])
    coarse = eulerSearch("C1", 15, 10, 15, 10)
    _, s0, _, best = collect(search_stage(al, coarse, support_frac=0.6))
    return dict(sc=sc, parts=parts, wedges=wedges, al=al, coarse=coarse, s0=s0, best=best)


def test_adjoint_against_central_differences_256(batch):
    al, coarse, s0, best = batch["al"], batch["coarse"], batch["s0"], batch["best"]
    g, fd = finite_difference_check(al, coarse[best], s0, h=1e-3, hs=1e-2)
    assert g.shape == (COUNT, 6) and np.all(np.isfinite(g))
    err_a = np.max(np.abs(g[:, :3] - fd[:, :3]), axis=1) / np.max(np.abs(fd[:, :3]), axis=1)
    err_s = np.max(np.abs(g[:, 3:] - fd[:, 3:]), axis=1) / np.max(np.abs(fd[:, 3:]), axis=1)
    print(f"adjoint vs central differences over {COUNT} particles: angles median {np.median(err_a):.2e} "
          f"max {np.max(err_a):.2e}; shifts median {np.median(err_s):.2e} max {np.max(err_s):.2e}")
    # the trilinear interpolant is piecewise linear: differences across cell
    # faces depart from the in-cell derivative; over the ~10^5 samples of a
    # 128^3 particle that averages to well under a percent
    assert np.median(err_a) < 0.01 and np.max(err_a) < 0.05, (np.argmax(err_a), np.max(err_a))
    assert np.median(err_s) < 0.01 and np.max(err_s) < 0.05, (np.argmax(err_s), np.max(err_s))


def test_adjoint_against_fp64_oracle_subset(batch):
    al, coarse, s0, best, parts, sc = (batch[k] for k in ("al", "coarse", "s0", "best", "parts", "sc"))
    cc, g = euler_eval(al, coarse[best], s0)
    F2 = O.batch_reference_spectrum(sc.ref, "c64")
    for i in ORACLE_SUBSET:
        R, dR = euler_derivatives(parts[i].start, *coarse[best[i]])
        bpp = O.batch_prepare(parts[i].data, sc.bandpass, batch["wedges"][i].get(), al.support)
        s_o, g_o = O.batch_eval(F2, N, bpp, R, s0[i], np.stack(dR))
        assert abs(cc[i] - s_o) < 1e-4 * abs(s_o), (i, cc[i], s_o)
        assert np.max(np.abs(g[i] - g_o)) < 2e-3 * np.max(np.abs(g_o)), (i, g[i], g_o)


def test_refinement_against_dense_grid(batch, ctx):
    parts, sc, coarse = batch["parts"][:DENSE_SUBSET], batch["sc"], batch["coarse"]
    al = BatchAligner(N, sc.ref, sc.bandpass, peak_radius=6, ctx=ctx)
    al.set_particles([p.data for p in parts], batch["wedges"][:DENSE_SUBSET], [p.start for p in parts])
    h = search_stage(al, coarse, support_frac=0.6)
    R0, s0 = h["R"].copy(), h["s"].copy()
    R, s, cc, _ = collect(refine_stage(h, 16, coarse=(0.6, 8)))
    Rj, sj, ccj, back = judge(al, R0.get(), s0.get(), R, s, cc)
    assert np.all(ccj >= cc) and np.all(ccj[back] > cc[back])
    dense = eulerSearch("C1", 15, 1, 15, 1)
    assert len(dense) == 23281
    Rd, sd, _, _ = collect(search_stage(al, dense, support_frac=1.0))
    acc_r = accuracy(parts, Rj, sj)
    acc_d = accuracy(parts, Rd, sd)
    print(f"refined from the 10-degree pick: {acc_r}; judge fallbacks {int(back.sum())}")
    print(f"dense 1-degree grid: {acc_d}")
    assert acc_r["within_1deg_0p5px"] >= acc_d["within_1deg_0p5px"]
    assert acc_r["median_angle_err_deg"] <= acc_d["median_angle_err_deg"]
    assert acc_r["fraction"] >= 0.9 and acc_d["fraction"] >= 0.85
