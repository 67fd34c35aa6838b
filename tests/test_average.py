"""Averaging on the CPU: the oracle's per-particle body pinned by exact grid
rotations (a 90 degree turn about z maps the grid onto itself), the
normalisation's closed form (mean 0, rms 1 inside the binary mask), linearity
in the weights, the half-set split and the skip of a flat particle."""
import math

import numpy as np

from oracle import emc_oracle as O


def _mask(n):
    z, y, x = np.meshgrid(*(np.arange(n) - n // 2,) * 3, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    return np.clip((n / 2 - 3 - r) / 2.0, 0.0, 1.0)


def test_exact_rotation_recovers_the_reference():
    rng = np.random.default_rng(0)
    n = 16
    ref = rng.standard_normal((n, n, n))
    M = O.define_matrix([90.0, 0.0, 0.0], "Bah", "inv")          # Rz(90): exact on the grid
    # the particle is the reference seen at orientation M: p(M u) = ref(u)
    part = O.xform3d(ref, M.T, [0, 0, 0], False)
    mask = _mask(n)
    odd, eve, _, _, used = O.average_particles([part], [M.reshape(-1, order="F")], [[0, 0, 0]], [1], [1.0], mask)
    b = mask > 0.01
    r = ref - ref[b].mean()
    r = r / math.sqrt(np.mean(r[b] ** 2)) * mask
    assert used.all() and np.allclose(odd, r, atol=1e-12) and not eve.any()


def test_normalisation_weights_halves_and_skip():
    rng = np.random.default_rng(1)
    n = 12
    mask = _mask(n)
    b = mask > 0.01
    wins = [rng.standard_normal((n, n, n)) + 3.0 for _ in range(3)] + [np.full((n, n, n), 2.0)]
    mats = [O.define_matrix(a, "Bah", "inv").reshape(-1, order="F") for a in ([10, 20, 30], [0, 0, 0],
                                                                             [200, 40, 5], [0, 0, 0])]
    sh = [[0.3, -0.2, 0.1], [0, 0, 0], [1, 0, 0], [0, 0, 0]]
    odd, eve, _, _, used = O.average_particles(wins[:1], mats[:1], sh[:1], [1], [1.0], mask)
    one = odd[b] / mask[b]
    assert abs(one.mean()) < 1e-12 and abs(np.sqrt(np.mean(one ** 2)) - 1.0) < 1e-12
    o2, _, _, _, _ = O.average_particles(wins[:1], mats[:1], sh[:1], [1], [2.5], mask)
    assert np.allclose(o2, 2.5 * odd, atol=1e-12)
    odd, eve, _, _, used = O.average_particles(wins, mats, sh, [1, 2, 2, 1], [1.0] * 4, mask)
    assert used.tolist() == [True, True, True, False]               # a flat particle has rms 0
    a, _, _, _, _ = O.average_particles(wins[:1], mats[:1], sh[:1], [1], [1.0], mask)
    _, e, _, _, _ = O.average_particles(wins[1:3], mats[1:3], sh[1:3], [2, 2], [1.0, 1.0], mask)
    assert np.allclose(odd, a, atol=1e-12) and np.allclose(eve, e, atol=1e-12)
