"""The plain single-subtomogram alignment at its workload size: a 64^3
reference of 40 Gaussian blobs at 2.5 A/px, the particle rotated and shifted
(U[-3, 3] px), weighted by a +-60 deg wedge and a 3 um CTF, SNR 0.1; the
coarse grid pinned to eulerSearch [180, 10, 10, 10] (1,236 orientations: every
out-of-plane direction at 10 deg, in-plane +-10 deg), bandpass 40-8 A.

Every orientation's CCC and shift through emc_align_grid (orientation blocks)
against the FP32 C restatement of the reference's body (oracle/emc_cpu.c,
itself pinned against the FP64 oracle): CCC to 1e-4 relative, shifts to
2e-3 voxel, the best orientation the same wherever the top-2 margin exceeds
twice the CCC tolerance; and the best grid orientation within the grid's
reach of the truth (under 10 deg)."""
import numpy as np
import pytest

from emclarity_amd import synthetic as S
from emclarity_amd.align_raw import ParticleAligner
from emclarity_amd.geometry import BH_defineMatrix, eulerSearch
from oracle import emc_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def cpu():
    try:
        from oracle import emc_cpu
        emc_cpu.load()
    except Exception as e:  # noqa: BLE001
        pytest.skip(f"oracle/lib/libemc_cpu.so not built ({e})")
    return emc_cpu


def test_config0_grid_matches_c_restatement(ctx, cpu):
    n = 64
    truth = (123.0, 47.0, 3.0)              # off-grid (phi, theta, psi): in reach of psi in [-10, 10]
    case = S.make_align_case(7, n=n, euler=truth, snr=0.1, n_blobs=40, blob_radius=0.3 * n)
    wedge = np.fft.fftshift(S.wedge_mask(n, 60.0)).astype(np.float32)
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5).astype(np.float32)
    pm = S.sphere_mask(n, n // 4)
    vm = S.sphere_mask(n, 0.4 * n)
    grid = np.ascontiguousarray(eulerSearch("C1", 180, 10, 10, 10))
    assert len(grid) == 1236
    al = ParticleAligner(case.particle, case.ref, wedge, bp, pm, vol_mask=vm)
    cc, est = al.grid(grid, np.zeros(3), mode=3)
    rc, re = cpu.align_grid(case.particle, case.ref, vm, pm, bp, wedge, np.eye(3), grid)
    assert np.max(np.abs(cc - rc) / np.abs(rc).max()) < 1e-4
    assert np.max(np.abs(est - re)) < 2e-3
    top = np.sort(rc)[-2:]
    if top[1] - top[0] > 2e-4 * abs(top[1]):
        assert int(np.argmax(cc)) == int(np.argmax(rc))
    b = int(np.argmax(cc))
    R = BH_defineMatrix([grid[b][0], grid[b][1], grid[b][2] - grid[b][0]], "Bah", "inv")
    ang = np.degrees(np.arccos(np.clip((np.trace(R.T @ case.rotmat) - 1) / 2, -1, 1)))
    print(f"config 0: best grid orientation {grid[b]} ({ang:.2f} deg from the truth), CCC {cc[b]:.4f}, "
          f"max |CCC - C| {np.max(np.abs(cc - rc)):.2e}, max |shift - C| {np.max(np.abs(est - re)):.2e}")
    assert ang < 10.0
