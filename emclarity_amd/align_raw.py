"""Per-particle alignment of BH_alignRaw3d_v2 on the GPU (the plain drop-in).

``align_particle`` follows alignment/BH_alignRaw3d_v2.m for one particle and
one reference: initial CCC (:1222-1232, :1355-1369), coarse grid (:1119-1419),
best-of with the initial pose (:1423-1445), local refinement (:1448-1644),
final translational pass (:1679-1765) and the 15 % CCC-drop revert
(:1785-1793). The orientation loops run inside emc_align_grid; this module
only does the host bookkeeping the MATLAB code does between them. The
azimuthal randomiser (:1142, drawn after rng('shuffle')) is fixed at zero so
results are reproducible.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import DeviceArray, as_device
from .fourier import FFTPlan


@dataclass
class AlignResult:
    row: np.ndarray          # [iRef, pIdx, phi, theta, psi-phi, CCC, 1, dx, dy, dz] (:1764-1765)
    ccc_initial: float
    coarse_ccc: np.ndarray
    coarse_shift: np.ndarray
    best_coarse: int
    refine_ccc: np.ndarray | None
    reverted: bool


class ParticleAligner:
    """Device-resident inputs of one particle and its references
    (the per-particle state of BH_alignRaw3d_v2.m:1084-1110)."""

    def __init__(self, particle, ref, wedge, bandpass, peak_mask, vol_mask=None, ref_wgt_rot=None,
                 ref_wgt=None, angles=None, shift_val=(0.0, 0.0, 0.0), peak_com=(3, 3, 3), pad_calc=None,
                 symmetry="C1", wCCC=None, ctx=None):
        """particle, ref, vol_mask: window W^3; wedge, bandpass, peak_mask,
        ref_wgt_rot, ref_wgt: calc C^3. pad_calc = [[low x,y,z], [top x,y,z]]
        (BH_multi_validArea's padCalc) with C = W + low + top; None: C = W.
        symmetry: the point group the particle and wedge rotations average
        over (C<n>, D<n>, O, I, I2; interpolator.m's table).

        Several references: ``ref`` a list of W^3 volumes (the half set's
        ref_FT2 of each), ``bandpass`` and ``ref_wgt`` one volume for all or
        a list, one per reference. Stage 1 rotates the first reference for
        every reference, as the reference code's refInterpolator holds
        ref_FT2_tmp{1}{1} (:1108); each reference gets its own ref_FT1,
        bandpass and weight in stage 2.

        wCCC: None (plain CCC, flgWeightCCC 0), a wCCC cell or ShellPlan for
        every reference, or a dict {reference index: cell or ShellPlan}: the
        shell-weighted CCC of BH_multi_xcf_Rotational in stage 2, shells over
        the calc grid."""
        from .geometry import symmetry_matrices
        mats = symmetry_matrices(symmetry)
        self.symmetry = symmetry
        self.nsym = len(mats)
        self.sym = np.ascontiguousarray(np.stack([m.reshape(-1, order="F") for m in mats]).ravel())
        self.particle = as_device(particle, np.float32, ctx)
        self.ctx = self.particle.ctx
        dims = self.particle.dims
        if not (dims[0] == dims[1] == dims[2]):
            raise ValueError("cubic boxes only")
        self.w = dims[0]
        pc = np.zeros((2, 3), int) if pad_calc is None else np.asarray(pad_calc, int).reshape(2, 3)
        calc = self.w + pc[0] + pc[1]
        if not (calc[0] == calc[1] == calc[2]) or pc.min() < 0:
            raise ValueError(f"padCalc {pc.tolist()} must pad the window to a cube")
        self.pad_calc = pc
        self.n = int(calc[0])
        dev = lambda a: None if a is None else as_device(a, np.float32, self.ctx)
        many = lambda a: isinstance(a, (list, tuple))
        refs = list(ref) if many(ref) else [ref]
        self.n_ref = len(refs)
        if self.n_ref < 1:
            raise ValueError("at least one reference")

        def per_ref(a, name):
            if not many(a):
                return [dev(a)] * self.n_ref
            if len(a) != self.n_ref:
                raise ValueError(f"{name}: {len(a)} entries for {self.n_ref} references")
            return [dev(x) for x in a]

        self.refs = [dev(r) for r in refs]
        self.ref = self.refs[0]
        self.bandpasses = per_ref(bandpass, "bandpass")
        self.ref_wgts = per_ref(ref_wgt, "ref_wgt")
        self.bandpass = self.bandpasses[0]
        self.ref_wgt = self.ref_wgts[0]
        self.wedge = dev(wedge)
        self.peak_mask = dev(peak_mask)
        self.vol_mask = dev(vol_mask)
        self.ref_wgt_rot = dev(ref_wgt_rot)
        if angles is None:
            self.angles = np.eye(3)
        else:
            a = np.asarray(angles, dtype=np.float64)
            # a geometry 9-vector (columns 17:25) is column-major, as reshape(angles, 3, 3)
            self.angles = a.reshape(3, 3) if a.ndim == 2 else a.reshape(3, 3, order="F")
        self.shift_val = np.asarray(shift_val, dtype=np.float64)
        self.peak_com = [int(c) for c in peak_com]
        calc_vols = [("wedge", self.wedge), ("peak_mask", self.peak_mask), ("ref_wgt_rot", self.ref_wgt_rot)]
        calc_vols += [("bandpass", x) for x in self.bandpasses] + [("ref_wgt", x) for x in self.ref_wgts]
        for name, a in calc_vols:
            if a is not None and a.dims != (self.n,) * 3:
                raise ValueError(f"{name} must be the calc size {self.n}^3, not {a.dims}")
        for name, a in [("ref", x) for x in self.refs] + [("vol_mask", self.vol_mask)]:
            if a is not None and a.dims != dims:
                raise ValueError(f"{name} must be the window size {self.w}^3, not {a.dims}")
        if self.bandpass is None or self.wedge is None or self.peak_mask is None:
            raise ValueError("bandpass, wedge and peak_mask are required")
        # ref_FT1 = conj(BH_bandLimitCenterNormalize(ref.*volMask, 1, ...)) (:607-608)
        self.plan = FFTPlan((self.n,) * 3, 1, self.ctx)
        self.ref_ft1s = []
        for r in self.refs:
            f = DeviceArray((self.n, self.n, self.n // 2 + 1), np.complex64, self.ctx)
            _lib.call("emc_ref_ft1", self.ctx.handle, self.plan.handle, r.vp,
                      self.vol_mask.vp if self.vol_mask is not None else None, _lib.i3(dims),
                      _lib.i3(pc[0]), _lib.i3(pc[1]), f.vp)
            self.ref_ft1s.append(f)
        self.ref_ft1 = self.ref_ft1s[0]
        from .xcf import ShellPlan
        plan = lambda c: None if c is None or (np.isscalar(c) and not c) else (
            c if isinstance(c, ShellPlan) else ShellPlan(c, self.ctx))
        if isinstance(wCCC, dict):
            self.shells = [plan(wCCC.get(r)) for r in range(self.n_ref)]
        else:
            one = plan(wCCC)
            self.shells = [one] * self.n_ref

    def grid_refs(self, angles_pts, est0, mode=3, refs=None):
        """emc_align_grid_padded over (phi, theta, psi) rows for the
        references ``refs`` (0-based, refToAlign; stage 1 uses the first one's
        bandpass) -> (ccc[n, len(refs)], estPeak[n, 3])."""
        refs = list(range(self.n_ref)) if refs is None else [int(r) for r in refs]
        if not refs or min(refs) < 0 or max(refs) >= self.n_ref:
            raise ValueError(f"reference indices {refs} out of 0..{self.n_ref - 1}")
        pts = np.ascontiguousarray(np.asarray(angles_pts, dtype=np.float64).reshape(-1, 3))
        n = pts.shape[0]
        nr = len(refs)
        ccc = np.zeros((n, nr))
        est = np.zeros((n, 3))
        dp = C.POINTER(C.c_double)
        vp = lambda a: a.vp if a is not None else None
        ptrs = lambda xs: (C.c_void_p * nr)(*[x.ptr if x is not None else None for x in xs])
        _lib.call("emc_align_grid_padded", self.ctx.handle, self.w, self.n, _lib.i3(self.pad_calc[0]),
                  self.nsym, _lib.d_arr(self.sym), nr, self.particle.vp, self.ref.vp,
                  vp(self.ref_wgt_rot), ptrs([self.ref_wgts[r] for r in refs]),
                  ptrs([self.ref_ft1s[r] for r in refs]),
                  (C.c_void_p * nr)(*[self.shells[r].handle.value if self.shells[r] is not None else None
                                      for r in refs]),
                  vp(self.vol_mask), self.peak_mask.vp,
                  ptrs([self.bandpasses[r] for r in refs]), self.wedge.vp,
                  _lib.d_arr(self.angles.reshape(-1, order="F")), _lib.d_arr(est0),
                  _lib.i3(self.peak_com), pts.ctypes.data_as(dp), n, int(mode),
                  ccc.ctypes.data_as(dp), est.ctypes.data_as(dp))
        return ccc, est

    def grid(self, angles_pts, est0, mode=3, ref=0):
        """One reference (index ``ref``): (ccc[n], estPeak[n, 3])."""
        ccc, est = self.grid_refs(angles_pts, est0, mode, [ref])
        return ccc[:, 0], est


def refine_list(best_phi, best_theta, best_psi, phi_inc, theta_inc, psi_inc):
    """Local grid around the coarse best (BH_alignRaw3d_v2.m:1457-1499); rows
    (phi, theta, psi) with the matrix built from [phi, theta, psi - phi]."""
    r_phi_inc = phi_inc / 4.0
    r_the_inc = theta_inc / 2.0
    r_psi_inc = psi_inc / 2.0
    psi_step = math.floor(math.sqrt(r_the_inc)) if r_the_inc else 1
    theta_step, phi_step = 1, 2

    def colon(a, s, b):
        if s == 0:
            return np.array([a])
        n = int(math.floor((b - a) / s + 1e-10)) + 1
        return a + s * np.arange(max(n, 0))

    if r_psi_inc == 0:
        in_plane = colon(best_psi - psi_step * r_the_inc / 2, r_the_inc / 2, best_psi + psi_step * r_the_inc / 2)
    else:
        in_plane = colon(best_psi - psi_step * r_psi_inc, r_psi_inc, best_psi + psi_step * r_psi_inc)
    polar = colon(best_theta - theta_step * r_the_inc, r_the_inc, best_theta + theta_step * r_the_inc)
    azim = colon(best_phi - phi_step * r_phi_inc, r_phi_inc, best_phi + phi_step * r_phi_inc)
    rows = [(p, t, s) for p in azim for t in polar for s in in_plane]
    return np.array(rows, dtype=np.float64)


def align_particle(al: ParticleAligner, grid_pts, phi_inc_of=None, theta_inc=0.0, psi_inc=0.0,
                   refine=True, i_ref=1, p_idx=1, refs=None) -> AlignResult:
    """One particle through the alignRaw stages (see module docstring).
    refs: 0-based references to align against (refToAlign, :1267-1276;
    default all of the aligner's). With one reference the row reports
    i_ref; with several, the 1-based index of the winning reference."""
    refs = list(range(al.n_ref)) if refs is None else [int(r) for r in refs]
    label = (lambda r: i_ref) if al.n_ref == 1 else (lambda r: r + 1)
    sv = al.shift_val
    ccc0, _ = al.grid_refs([[0.0, 0.0, 0.0]], sv, mode=2, refs=refs)   # initial CCC per reference
    k0 = int(np.flatnonzero(ccc0[0] == ccc0[0].max())[0])             # :1425-1430
    ccc_initial = float(ccc0[0, k0])
    ref_initial = refs[k0]
    cc2, est = al.grid_refs(grid_pts, sv, mode=3, refs=refs)          # coarse, rows orientation-major
    flat = cc2.ravel()
    bi = int(np.flatnonzero(flat == flat.max())[0])                   # sortrows(-6), stable
    b, br = divmod(bi, len(refs))
    ccc = cc2[:, br]
    best = dict(phi=grid_pts[b][0], theta=grid_pts[b][1], psi=grid_pts[b][2], ccc=flat[bi], est=est[b],
                ref=refs[br])
    if ccc_initial > best["ccc"]:                             # :1433-1435
        best = dict(phi=0.0, theta=0.0, psi=0.0, ccc=ccc_initial, est=sv.copy(), ref=ref_initial)
    refine_ccc = None
    row_angles = (best["phi"], best["theta"], best["psi"] - best["phi"])
    est_best = best["est"]
    ccc_best = best["ccc"]
    if refine:
        phi_inc = phi_inc_of(best["theta"]) if callable(phi_inc_of) else (phi_inc_of or 0.0)
        rl = refine_list(best["phi"], best["theta"], best["psi"], phi_inc, theta_inc, psi_inc)
        rc, re = al.grid(rl, best["est"], mode=3, ref=best["ref"])        # rRef only (:1560-1640)
        refine_ccc = rc
        k = int(np.flatnonzero(rc == rc.max())[0])
        row_angles = (rl[k][0], rl[k][1], rl[k][2] - rl[k][0])
        est_best = re[k]
        ccc_best = rc[k]
    # final translational pass at the best pose (:1679-1765): stage 1 only,
    # matrix from [phi, theta, psi'] with psi' already psi - phi
    fin = np.array([[row_angles[0], row_angles[1], row_angles[2] + row_angles[0]]])
    _, fest = al.grid(fin, est_best, mode=1, ref=best["ref"])          # finalRef (:1679)
    shift = fest[0] - sv                                       # :1764-1765
    row = np.array([label(best["ref"]), p_idx, *row_angles, ccc_best, 1.0, *shift])
    reverted = False
    d = row[5] - ccc_initial
    if d < 0 and abs(d) > 0.15 * ccc_initial:                  # :1785-1793
        row = np.array([label(ref_initial), p_idx, 0.0, 0.0, 0.0, ccc_initial, 1.0, *sv])
        reverted = True
    return AlignResult(row, ccc_initial, ccc, est, b, refine_ccc, reverted)
