"""GPU parity of the batched search and the adjoint at the box sizes and
launch shapes of the mixed-size alignment workload (96^3 and 128^3, the
shipped MI355X tuning table applied: its search patches, passes per launch
and adjoint workgroup counts), for both reference-table formats, against the
FP64 oracle on the full frequency grid.

Search: every (particle, orientation) score to 1e-4 relative and its
sub-pixel shift to 2e-3 px, over the whole support and over 0.6 x the
support radius (the coarse search's setting); the best orientation exact
wherever the oracle's top-2 margin exceeds twice the score tolerance.
Adjoint: the score to 1e-4 relative and the 6-component gradient to 2e-3 of
its largest component."""
import numpy as np
import pytest

from emclarity_amd.autotune import params_for
from emclarity_amd.batched import BatchAligner, euler_derivatives
from emclarity_amd.masking import BH_weightMaskMex
from emclarity_amd.workload import local_matrix, make_particle, make_size_class
from oracle import emc_oracle as O

pytestmark = pytest.mark.gpu

SCORE_RTOL = 1e-4
SHIFT_TOL = 2e-3


def _particles(n, count, seed):
    rng = np.random.default_rng(seed)
    sc, blobs = make_size_class(n)
    parts = [make_particle(rng, sc, blobs) for _ in range(count)]
    wedges = [BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]]).get()
              for p in parts]
    return sc, parts, wedges


def _orientations(p):
    """8 local orientations around the start: the off-grid truth, points of
    the 10-degree coarse grid and a few far ones."""
    tl = p.truth_local
    return np.array([tl, (0.0, 0.0, 0.0), (0.0, 10.0, -10.0), (90.0, 10.0, 0.0), (180.0, 10.0, 10.0),
                     (tl[0] + 2.0, tl[1] + 1.0, tl[2] - 1.0), (270.0, 10.0, 10.0), (45.0, 5.0, -5.0)])


@pytest.mark.parametrize("table", ["c64", "f16"])
@pytest.mark.parametrize("n", [96, 128])
def test_batched_search_at_workload_sizes(ctx, n, table):
    tune = params_for(n, table)
    assert tune, f"no shipped tuning for {n} {table}"
    sc, parts, wedges = _particles(n, 2, 300 + n)
    al = BatchAligner(n, sc.ref, sc.bandpass, peak_radius=6, table=table, ctx=ctx)
    assert al._tune == tune
    al.set_particles([p.data for p in parts], wedges, [p.start for p in parts])
    F2 = O.batch_reference_spectrum(sc.ref, table)
    for frac in (1.0, 0.6):
        radius = frac * al.support
        for pi, p in enumerate(parts):
            pts = _orientations(p)
            # one orientation set per particle: search both particles with
            # this particle's set and check only its row
            score, shift = al.search(pts, support_radius=radius)
            bpp = O.batch_prepare(p.data, sc.bandpass, wedges[pi], radius)
            ref = []
            for k, q in enumerate(pts):
                R = p.start @ local_matrix(*q)
                s_o, sh_o = O.batch_score_map(F2, n, bpp, R, 6)
                ref.append(s_o)
                assert abs(score[pi, k] - s_o) < SCORE_RTOL * abs(s_o), (frac, pi, k, score[pi, k], s_o)
                assert np.max(np.abs(shift[pi, k] - sh_o)) < SHIFT_TOL, (frac, pi, k, shift[pi, k], sh_o)
            ref = np.array(ref)
            top = np.sort(ref)[-2:]
            if top[1] - top[0] > 2 * SCORE_RTOL * abs(top[1]):
                assert int(np.argmax(score[pi])) == int(np.argmax(ref)), (frac, pi, score[pi], ref)


@pytest.mark.parametrize("table", ["c64", "f16"])
@pytest.mark.parametrize("n", [96, 128])
def test_adjoint_at_workload_sizes(ctx, n, table):
    sc, parts, wedges = _particles(n, 3, 400 + n)
    al = BatchAligner(n, sc.ref, sc.bandpass, peak_radius=6, table=table, ctx=ctx)
    al.set_particles([p.data for p in parts], wedges, [p.start for p in parts])
    F2 = O.batch_reference_spectrum(sc.ref, table)
    Rs, dRs, ss = [], [], []
    for p in parts:
        phi, theta, psi = p.truth_local
        R, dR = euler_derivatives(p.start, phi + 1.3, theta - 0.7, psi + 0.9)
        Rs.append(R)
        dRs.append(np.stack(dR))
        ss.append(p.shift + np.array([0.3, -0.2, 0.25]))
    cc, g = al.eval_grad(np.stack(Rs), np.stack(dRs), np.stack(ss))
    for i, p in enumerate(parts):
        bpp = O.batch_prepare(p.data, sc.bandpass, wedges[i], al.support)
        s_o, g_o = O.batch_eval(F2, n, bpp, Rs[i], ss[i], dRs[i])
        assert abs(cc[i] - s_o) < SCORE_RTOL * abs(s_o), (i, cc[i], s_o)
        assert np.max(np.abs(g[i] - g_o)) < 2e-3 * np.max(np.abs(g_o)), (i, g[i], g_o)
