"""The adjoint of the rotate-weight-correlate chain behind a
torch.autograd.Function, for gradcheck only.

``AlignScore.apply(angles, shifts, aligner)`` returns each particle's
normalised CC at the pose R_p = start_p Rz(phi) Rx(theta) Rz(psi - phi)
(BH_defineMatrix([phi, theta, psi - phi], 'Bah', 'inv') right of the
particle's start matrix; angles in radians) and shift s_p (pixels); its
backward multiplies the incoming gradient by the adjoint's d CC / d(phi,
theta, psi, s) from ``emc_align_eval_grad``. The tensors stay on the CPU in
float64: the library owns the device work, and torch's own GPU runtime is
never initialised (the library is loaded before torch is imported, so its
kernels run on the HIP runtime they were built against).

``python -m emclarity_amd.autograd`` runs torch.autograd.gradcheck on a
small batch and prints one line; it exits non-zero on failure.
"""
from __future__ import annotations

import sys

import numpy as np

from . import _lib


def _torch():
    _lib.load()            # the package's HIP runtime first
    import torch
    return torch


def align_score_function():
    """The autograd.Function class (built on first use, importing torch)."""
    torch = _torch()
    from .batched import euler_derivatives

    class AlignScore(torch.autograd.Function):
        @staticmethod
        def forward(ctx, angles, shifts, aligner):
            a = angles.detach().cpu().numpy().astype(np.float64)
            s = shifts.detach().cpu().numpy().astype(np.float64)
            P = a.shape[0]
            if P != aligner.n_part or a.shape != (P, 3) or s.shape != (P, 3):
                raise ValueError("angles and shifts must be (particles, 3)")
            R = np.empty((P, 3, 3))
            dR = np.empty((P, 3, 3, 3))
            for p in range(P):
                Rp, dRp = euler_derivatives(aligner.starts[p], *np.degrees(a[p]))
                R[p], dR[p] = Rp, np.stack(dRp)
            cc, g = aligner.eval_grad(R, dR, s)
            ctx.save_for_backward(torch.from_numpy(np.asarray(g, dtype=np.float64)))
            return torch.from_numpy(np.asarray(cc, dtype=np.float64))

        @staticmethod
        def backward(ctx, grad_cc):
            (g,) = ctx.saved_tensors
            gc = grad_cc.reshape(-1, 1).to(g.dtype)
            return gc * g[:, :3], gc * g[:, 3:], None

    return AlignScore


def gradcheck_demo(n=32, table="c64", seed=11):
    """torch.autograd.gradcheck of AlignScore on two particles with central
    differences h = 1e-3 rad / 1e-3 px. The trilinear interpolant is
    piecewise linear, so its in-cell derivative and a difference across cell
    faces differ by a few percent: the tolerance is 5 % of the largest
    gradient component."""
    torch = _torch()
    from . import synthetic as S
    from .batched import BatchAligner
    from .masking import BH_bandpass3d, BH_weightMaskMex
    from .workload import local_matrix
    AlignScore = align_score_function()
    c = S.make_align_case(seed, n=n, euler=(40.0, 30.0, 20.0), snr=1.0, n_blobs=16, blob_radius=0.22 * n)
    w = BH_weightMaskMex((n, n, n), 1, S.tilt_table(60.0, 3.0, 3.0, 2.5), [n // 2 + 1] * 3,
                         [[n, n, n], [0, 0, 0]]).get()
    bp = BH_bandpass3d((n, n, n), 1e-3, 40.0, 8.0, "GPU", 2.5).get()
    al = BatchAligner(n, c.ref, bp, peak_radius=6, table=table)
    al.set_particles([c.particle, c.particle], [w, w], [np.eye(3), local_matrix(20.0, 10.0, 5.0)])
    angles = torch.tensor(np.radians([[43.0, 27.5, 22.0], [20.0, 25.0, 14.0]]), dtype=torch.float64,
                          requires_grad=True)
    shifts = torch.tensor([[-1.2, -2.0, -0.4], [0.3, 0.5, -0.8]], dtype=torch.float64, requires_grad=True)
    fn = lambda a, s: AlignScore.apply(a, s, al)
    out = fn(angles, shifts)
    g_a, g_s = torch.autograd.grad(out.sum(), (angles, shifts))
    scale = float(max(g_a.abs().max(), g_s.abs().max()))
    ok = torch.autograd.gradcheck(fn, (angles, shifts), eps=1e-3, atol=0.05 * scale, rtol=0.05,
                                  raise_exception=False)
    return bool(ok), out.detach().numpy(), scale


def main():
    ok, cc, scale = gradcheck_demo()
    print(f"gradcheck {'passed' if ok else 'FAILED'}: cc {np.round(cc, 4).tolist()}, max |grad| {scale:.4f}",
          flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
