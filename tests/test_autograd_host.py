"""Host side of the autograd wrapper: its pose parameterisation is the
reference's BH_defineMatrix([phi, theta, psi - phi], 'Bah', 'inv') right of
the start matrix, and its matrix derivatives match central differences."""
import numpy as np

from emclarity_amd.batched import euler_derivatives
from emclarity_amd.geometry import BH_defineMatrix


def test_pose_parameterisation_and_derivatives():
    rng = np.random.default_rng(3)
    A = BH_defineMatrix(rng.uniform(-180, 180, 3), "Bah", "inv")
    ang = (43.0, 27.5, 22.0)
    R, dR = euler_derivatives(A, *ang)
    assert np.allclose(R, A @ BH_defineMatrix([ang[0], ang[1], ang[2] - ang[0]], "Bah", "inv"), atol=1e-12)
    h = 1e-6
    for i in range(3):
        ap, am = list(ang), list(ang)
        ap[i] += np.degrees(h)
        am[i] -= np.degrees(h)
        fd = (euler_derivatives(A, *ap)[0] - euler_derivatives(A, *am)[0]) / (2 * h)
        assert np.allclose(dR[i], fd, atol=1e-8), i


def test_autograd_function_builds():
    from emclarity_amd.autograd import align_score_function
    F = align_score_function()
    assert hasattr(F, "apply") and hasattr(F, "backward")
