"""Alignment outputs on the CPU: the Euler decomposition behind the rows
against BH_defineMatrix (oracle), the resume-file text, the row selection
and CCC revert, and BH_rawAlignmentsApply (host mirror, vectorised) against
the particle-by-particle oracle, whose own pins are hand-worked compositions."""
import math

import numpy as np
import pytest

from emclarity_amd import _lib
from emclarity_amd.alignment_outputs import (BH_rawAlignmentsApply, bah_angles_from_matrix, format_align_resume,
                                             read_align_resume, revert_drops, unique_particle_rows,
                                             write_align_resume)
from oracle import emc_oracle as O


def _lib_ok():
    try:
        _lib.load()
        return True
    except Exception:
        return False


needs_lib = pytest.mark.skipif(not _lib_ok(), reason="libemc_hip.so not built")


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def test_euler_decomposition_round_trip():
    rng = np.random.default_rng(0)
    Ms = [_random_rotation(rng) for _ in range(200)]
    Ms += [O.define_matrix([a, t, g], "Bah", "inv") for a, t, g in
           [(30, 0, 0), (30, 0, 40), (-70, 180, 0), (10, 180, 25), (0, 90, 0), (200, 45, -300)]]
    A = bah_angles_from_matrix(np.stack(Ms))
    for M, a in zip(Ms, A):
        assert 0.0 <= a[1] <= 180.0
        assert np.max(np.abs(O.define_matrix(a, "Bah", "inv") - M)) < 1e-12
    # a grid point comes back as itself (theta in (0, 180), angles in (-180, 180])
    a = bah_angles_from_matrix(O.define_matrix([40.0, 9.0, -25.0], "Bah", "inv"))
    assert np.allclose(a, [40.0, 9.0, -25.0])


def test_resume_text_and_reading(tmp_path):
    rows = np.array([[1, 17, 12.3456, 3.0, -0.5, 0.123456789, 1, 0.25, -1.125, 2.0],
                     [2, 18, -170.0, 179.9994, 5.0, 0.5, 1, 0, 0, 0],
                     [1.5, 19, 0, 0, 0, -0.25, 1, 1e-4, -2e-4, 3]])
    txt = format_align_resume(rows)
    assert txt == O.align_resume_text(rows)
    assert txt.splitlines()[0] == "1 17 12.346  3.000 -0.500 0.123457 1.000000  0.250 -1.125  2.000 "
    assert txt.splitlines()[2].startswith("1.500000e+00 19 ")
    p = tmp_path / "alignResume" / "run" / "tomo_1.txt"
    write_align_resume(p, rows)
    back = read_align_resume(p)
    assert back.shape == (3, 10) and np.allclose(back, rows, atol=5e-4)
    two = np.concatenate([rows, rows + 1], axis=1)
    assert format_align_resume(two, 2) == O.align_resume_text(two, 2)


def test_row_selection_and_revert():
    s = np.array([[1, 5, 0, 0, 0, .4, 1, 0, 0, 0], [1, 0, 0, 0, 0, .9, 1, 0, 0, 0],
                  [1, 3, 0, 0, 0, .5, 1, 0, 0, 0], [1, 5, 9, 9, 9, .8, 1, 0, 0, 0],
                  [1, 2, 0, 0, 0, .1, 1, 0, 0, 0]], float)
    got = unique_particle_rows(s)
    assert np.array_equal(got, O.unique_particle_rows(s))
    assert got[:, 1].tolist() == [5, 3, 2] and got[0, 5] == .4
    best = np.array([[1, 1, 5, 5, 5, 0.80, 1, 1, 1, 1], [1, 2, 5, 5, 5, 0.50, 1, 1, 1, 1],
                     [1, 3, 5, 5, 5, 0.88, 1, 1, 1, 1]])
    init = np.array([[1, 1, 0, 0, 0, 0.70, 1, 0, 0, 0], [1, 2, 0, 0, 0, 0.60, 1, 0, 0, 0],
                     [1, 3, 0, 0, 0, 1.00, 1, 0, 0, 0]])
    out = revert_drops(best, init)
    # 0.50 vs 0.60: drop 0.10 > 0.09 -> initial; 0.88 vs 1.00: drop 0.12 < 0.15 -> kept
    assert np.array_equal(out[0], best[0]) and np.array_equal(out[1], init[1]) and np.array_equal(out[2], best[2])


def _geometry(rng, n, n_peaks=1):
    g = np.zeros((n, 26 * n_peaks))
    for p in range(n_peaks):
        o = 26 * p
        g[:, 0 + o] = rng.uniform(0, 1, n)
        g[:, 1 + o] = 2
        g[:, 3 + o] = np.arange(1, n + 1)        # the same particle id for every peak
        g[:, 10 + o:13 + o] = rng.uniform(50, 400, (n, 3))
        g[:, 16 + o:25 + o] = np.stack([_random_rotation(rng).reshape(-1, order="F") for _ in range(n)])
        g[:, 25 + o] = 1
    return g


def test_oracle_apply_pins():
    rng = np.random.default_rng(1)
    g = _geometry(rng, 3)
    rows = np.array([[4, 1, 0, 0, 0, 0.7, 0.3, 0, 0, 0],
                     [4, 2, 90, 0, 0, 0.6, 0.3, 1, -2, 0.5],
                     [4, 3, 0, 0, 0, 0.5, 0.3, 0, 0, 0]])
    g[2, 25] = -9999
    out = O.raw_alignments_apply({"t": g}, {"t": rows}, 2, 1, update_weights=True, update_class=True)["t"]
    # zero angles and shifts: orientation and position unchanged, CCC/weight/class updated
    assert np.allclose(out[0, 16:25], g[0, 16:25]) and np.allclose(out[0, 10:13], g[0, 10:13])
    assert (out[0, 0], out[0, 1], out[0, 25]) == (0.7, 0.3, 4)
    # (90, 0, 0) right-multiplies Rz(90); shifts scale by the sampling
    old = g[1, 16:25].reshape(3, 3, order="F")
    rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    assert np.allclose(out[1, 16:25].reshape(3, 3, order="F"), old @ rz)
    assert np.allclose(out[1, 10:13], g[1, 10:13] + [2, -4, 1])
    # ignored particles stay untouched
    assert np.array_equal(out[2], g[2])
    # a particle without a row is ignored from now on
    out = O.raw_alignments_apply({"t": g}, {"t": rows[1:2]}, 1, 1)["t"]
    assert out[0, 25] == -9999 and out[1, 25] == 1


@needs_lib
@pytest.mark.parametrize("n_peaks", [1, 2])
def test_apply_matches_oracle(n_peaks):
    rng = np.random.default_rng(10 + n_peaks)
    geom = {"tomo_1": _geometry(rng, 40, n_peaks), "tomo_2": _geometry(rng, 15, n_peaks)}
    geom["tomo_1"][[3, 7], 25] = -9999
    best = {}
    for t, g in geom.items():
        keep = rng.permutation(g.shape[0])[: g.shape[0] - 4]
        r = np.zeros((keep.size, 10 * n_peaks))
        for p in range(n_peaks):
            q = 10 * p
            r[:, q] = rng.integers(1, 4, keep.size)
            r[:, q + 1] = g[keep, 3]           # peak 1 ids (the rows' id column is per peak 1)
            r[:, q + 2:q + 5] = rng.uniform(-20, 20, (keep.size, 3))
            r[:, q + 5] = rng.uniform(0, 1, keep.size)
            r[:, q + 6] = rng.uniform(0, 1, keep.size)
            r[:, q + 7:q + 10] = rng.uniform(-3, 3, (keep.size, 3))
        best[t] = r
    for uw, uc in [(False, False), (True, True)]:
        got = BH_rawAlignmentsApply(geom, best, 2, n_peaks, "Bah", uw, uc)
        ref = O.raw_alignments_apply(geom, best, 2, n_peaks, "Bah", uw, uc)
        for t in geom:
            assert np.allclose(got[t], ref[t], atol=1e-12, rtol=0), t
