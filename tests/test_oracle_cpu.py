"""The C restatement of the plain alignment body (oracle/emc_cpu.c, FP32,
OpenMP) against the FP64 numpy oracle: CCC to 1e-4 relative, shifts to 2e-3
voxel, the same best orientation, any thread count."""
import numpy as np
import pytest

from emclarity_amd import synthetic as S
from oracle import emc_oracle as O


@pytest.fixture
def cpu():
    """The C restatement, loaded only when a test of this file runs (a GPU
    session that deselects these tests never maps the library)."""
    try:
        from oracle import emc_cpu
        emc_cpu.load()
    except Exception as e:  # noqa: BLE001
        pytest.skip(f"oracle/lib/libemc_cpu.so not built ({e})")
    return emc_cpu


@pytest.mark.parametrize("n", [24, 32])
def test_c_restatement_matches_numpy_oracle(cpu, n):
    truth = (40.0, 30.0, 20.0)
    case = S.make_align_case(3, n=n, euler=truth, snr=2.0, n_blobs=16, blob_radius=0.22 * n)
    wedge = np.fft.fftshift(S.wedge_mask(n, 60.0)).astype(np.float32)
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5).astype(np.float32)
    pm = S.sphere_mask(n, n // 4)
    vm = S.sphere_mask(n, 0.4 * n)
    pts = np.array([truth, (0.0, 0.0, 0.0), (80.0, 30.0, 20.0), (10.0, 5.0, -5.0)])
    inp = O.AlignInputs(particle=case.particle.astype(np.float64), ref=case.ref.astype(np.float64),
                        ref_wgt_rot=None, ref_wgt=1.0,
                        ref_ft1=O.ref_ft1_of(case.ref.astype(np.float64), vm, [[0] * 3, [0] * 3]),
                        vol_mask=vm.astype(np.float64), peak_mask=pm.astype(np.float64),
                        bandpass=bp.astype(np.float64), wedge=wedge.astype(np.float64),
                        angles=np.eye(3), shift_val=np.zeros(3), pad_calc=[[0] * 3, [0] * 3])
    rc, re = O.align_grid(inp, pts)
    for threads in (1, 3):
        cc, est = cpu.align_grid(case.particle, case.ref, vm, pm, bp, wedge, np.eye(3), pts, threads=threads)
        assert np.max(np.abs(cc - rc) / np.abs(rc)) < 1e-4
        assert np.max(np.abs(est - re)) < 2e-3
        assert int(np.argmax(cc)) == int(np.argmax(rc)) == 0
