"""GPU parity of the batched Fourier-space alignment and of its adjoint
against the FP64 oracle (full-grid restatement), finite differences, and
accuracy against synthetic ground truth."""
import numpy as np
import pytest

from emclarity_amd import synthetic as S
from emclarity_amd.batched import BatchAligner, align_batch, euler_derivatives, rodrigues
from emclarity_amd.geometry import eulerSearch
from emclarity_amd.masking import BH_weightMaskMex
from emclarity_amd.workload import angle_error_deg, local_matrix
from oracle import emc_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("table", ["c64", "f16"])
def test_batched_search_matches_oracle(ctx, table):
    n = 32
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5)
    truth = (40.0, 30.0, 20.0)
    c = S.make_align_case(11, n=n, euler=truth, snr=1.0, n_blobs=16, blob_radius=0.22 * n)
    T = S.tilt_table(60.0, 3.0, 3.0, 2.5)
    w = BH_weightMaskMex((n, n, n), 1, T, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]]).get()
    # second particle: the same data seen from a different start orientation
    A2 = local_matrix(20.0, 10.0, 5.0)
    al = BatchAligner(n, c.ref, bp.astype(np.float32), peak_radius=6, table=table)
    al.set_particles([c.particle, c.particle], [w, w], [np.eye(3), A2])
    pts = np.array([truth, (0.0, 0.0, 0.0), (40.0, 30.0, 30.0), (80.0, 30.0, 20.0), (200.0, 60.0, -40.0),
                    (45.0, 32.0, 18.0), (10.0, 45.0, 120.0)])
    score, shift = al.search(pts)
    F2 = O.batch_reference_spectrum(c.ref, table)
    bpp = O.batch_prepare(c.particle, bp, w, al.support)
    for pi, A in enumerate([np.eye(3), A2]):
        rs = []
        for k, p in enumerate(pts):
            R = A @ local_matrix(*p)
            sc, sh = O.batch_score_map(F2, n, bpp, R, 6)
            rs.append(sc)
            assert abs(score[pi, k] - sc) < 1e-4 * abs(sc), (pi, k, score[pi, k], sc)
            assert np.max(np.abs(shift[pi, k] - sh)) < 2e-3, (pi, k, shift[pi, k], sh)
        rs = np.array(rs)
        top2 = np.sort(rs)[-2:]
        if top2[1] - top2[0] > 1e-3 * top2[1]:
            assert int(np.argmax(score[pi])) == int(np.argmax(rs))


@pytest.mark.parametrize("table", ["c64", "f16"])
def test_adjoint_matches_oracle_and_finite_differences(ctx, table):
    n = 32
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5)
    c = S.make_align_case(11, n=n, euler=(40.0, 30.0, 20.0), snr=1.0, n_blobs=16, blob_radius=0.22 * n)
    w = BH_weightMaskMex((n, n, n), 1, S.tilt_table(60.0, 3.0, 3.0, 2.5), [n // 2 + 1] * 3,
                         [[n, n, n], [0, 0, 0]]).get()
    al = BatchAligner(n, c.ref, bp.astype(np.float32), peak_radius=6, table=table)
    al.set_particles([c.particle], [w], [np.eye(3)])
    F2 = O.batch_reference_spectrum(c.ref, table)
    bpp = O.batch_prepare(c.particle, bp, w, al.support)
    # Euler-angle gradient at an off-grid pose
    ang = (43.0, 27.5, 22.0)
    s = np.array([-1.2, -2.0, -0.4])
    R, dR = euler_derivatives(np.eye(3), *ang)
    cc, g = al.eval_grad(R[None], np.stack(dR)[None], s[None])
    ocs, og = O.batch_eval(F2, n, bpp, R, s, dR)
    assert abs(cc[0] - ocs) < 1e-4 * abs(ocs)
    assert np.max(np.abs(g[0] - og)) < 2e-3 * np.max(np.abs(og)), (g[0], og)
    # central differences of the device score, h = 1e-3 rad / 1e-2 px
    h, hs = 1e-3, 1e-2
    fd = []
    for i in range(3):
        a_p, a_m = list(ang), list(ang)
        a_p[i] += np.degrees(h)
        a_m[i] -= np.degrees(h)
        Rp, _ = euler_derivatives(np.eye(3), *a_p)
        Rm, _ = euler_derivatives(np.eye(3), *a_m)
        cp, _ = al.eval_grad(Rp[None], np.stack(dR)[None], s[None])
        cm, _ = al.eval_grad(Rm[None], np.stack(dR)[None], s[None])
        fd.append((cp[0] - cm[0]) / (2 * h))
    for i in range(3):
        e = np.zeros(3)
        e[i] = hs
        cp, _ = al.eval_grad(R[None], np.stack(dR)[None], (s + e)[None])
        cm, _ = al.eval_grad(R[None], np.stack(dR)[None], (s - e)[None])
        fd.append((cp[0] - cm[0]) / (2 * hs))
    fd = np.array(fd)
    # the trilinear interpolant is piecewise linear: FD across cell faces
    # differs from the in-cell derivative by a few percent
    assert np.max(np.abs(g[0] - fd)) < 0.05 * np.max(np.abs(fd)), (g[0], fd)


@pytest.mark.parametrize("table", ["c64", "f16"])
def test_batched_refinement_accuracy(ctx, table):
    """64^3 particles with their own wedge and defocus, SNR 0.1: coarse
    6-degree local grid + adjoint refinement reaches <=1 deg and <=0.5 px."""
    n = 64
    rng = np.random.default_rng(9)
    from emclarity_amd.workload import make_particle, make_size_class
    sc, blobs = make_size_class(n)
    parts = [make_particle(rng, sc, blobs) for _ in range(4)]
    wedges = [BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]]).get()
              for p in parts]
    al = BatchAligner(n, sc.ref, sc.bandpass, peak_radius=6, table=table)
    al.set_particles([p.data for p in parts], wedges, [p.start for p in parts])
    grid = eulerSearch("C1", 15, 6, 15, 6)
    R, s, cc, _ = align_batch(al, grid, refine_iters=12)
    ok = 0
    for i, p in enumerate(parts):
        ea = angle_error_deg(R[i], p.truth_R)
        es = float(np.max(np.abs(s[i] - p.shift)))
        ok += int(ea <= 1.0 and es <= 0.5)
    assert ok >= 3, [(angle_error_deg(R[i], p.truth_R), s[i] - p.shift) for i, p in enumerate(parts)]


def test_coarse_to_fine_refinement_accuracy(ctx):
    """The bench's schedule: the 10-degree coarse grid, then 8 iterations
    over 0.6 x the support radius and 8 over the full support
    (refine_stage(coarse=...)) reaches <=1 deg and <=0.5 px; the first stage
    alone moves every particle off its grid pick."""
    from emclarity_amd.batched import collect, refine_stage, search_stage
    from emclarity_amd.workload import make_particle, make_size_class
    n = 64
    rng = np.random.default_rng(19)
    sc, blobs = make_size_class(n)
    parts = [make_particle(rng, sc, blobs) for _ in range(6)]
    wedges = [BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]]).get()
              for p in parts]
    al = BatchAligner(n, sc.ref, sc.bandpass, peak_radius=6)
    al.set_particles([p.data for p in parts], wedges, [p.start for p in parts])
    grid = eulerSearch("C1", 15, 10, 15, 10)
    h0 = search_stage(al, grid)
    R0 = h0["R"].get().reshape(-1, 3, 3)
    R, s, cc, _ = collect(refine_stage(search_stage(al, grid), 16, coarse=(0.6, 8)))
    ok = sum(int(angle_error_deg(R[i], p.truth_R) <= 1.0 and np.max(np.abs(s[i] - p.shift)) <= 0.5)
             for i, p in enumerate(parts))
    assert ok >= 5, [(angle_error_deg(R[i], p.truth_R), s[i] - p.shift) for i, p in enumerate(parts)]
    Rc, _, _, _ = collect(refine_stage(search_stage(al, grid), 8, coarse=(0.6, 8)))
    assert all(angle_error_deg(Rc[i], R0[i]) > 1e-3 for i in range(len(parts)))


def test_batched_preparation_matches_single(ctx):
    """emc_batch_prepare_particles (chunks 16 + 2 + 1, some without wedge)
    against emc_batch_prepare_particle per particle, and A against the
    oracle's preparation."""
    import ctypes as C
    from emclarity_amd import _lib
    from emclarity_amd.device import DeviceArray
    n, P = 16, 19
    rng = np.random.default_rng(5)
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5)
    w = BH_weightMaskMex((n, n, n), 1, S.tilt_table(60.0, 3.0, 3.0, 2.5), [n // 2 + 1] * 3,
                         [[n, n, n], [0, 0, 0]]).get()
    vols = [rng.standard_normal((n, n, n)).astype(np.float32) for _ in range(P)]
    wedges = [None if i % 5 == 3 else w for i in range(P)]
    al = BatchAligner(n, np.zeros((n, n, n), np.float32), bp.astype(np.float32))
    al.set_particles(vols, wedges, [np.eye(3)] * P)
    Qb, Nb, Ab = al.Q.get(), al.N.get(), al.A.get()
    nh = al.nh
    d_bp = DeviceArray.from_numpy(bp.astype(np.float32))
    for i in range(P):
        Q1 = DeviceArray((nh,), np.complex64)
        N1 = DeviceArray((nh,), np.float32)
        A1 = DeviceArray((1,), np.float64)
        pd = DeviceArray.from_numpy(vols[i])
        wd = None if wedges[i] is None else DeviceArray.from_numpy(wedges[i].astype(np.float32))
        _lib.call("emc_batch_prepare_particle", ctx.handle, n, pd.vp, d_bp.vp,
                  None if wd is None else wd.vp, Q1.vp, N1.vp, A1.vp)
        q1, n1, a1 = Q1.get(), N1.get(), A1.get()[0]
        scale = np.max(np.abs(q1))
        assert np.max(np.abs(Qb[i] - q1)) <= 1e-5 * scale, i
        assert np.array_equal(Nb[i], n1), i
        assert abs(Ab[i] - a1) <= 1e-6 * a1, (i, Ab[i], a1)
        ref = O.batch_prepare(vols[i].astype(np.float64), bp, wedges[i], al.support)
        assert abs(Ab[i] - ref.A) <= 1e-5 * ref.A, (i, Ab[i], ref.A)


def test_masked_preparation_equals_premultiplied_particles(ctx):
    """emc_batch_prepare_particles_masked (19 particles: chunks 16 + 2 + 1)
    equals the unmasked preparation of the particles multiplied by the mask
    on the host, bit for bit (one FP32 product per voxel either way); a mask
    of ones equals no mask."""
    n, P = 16, 19
    rng = np.random.default_rng(8)
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5).astype(np.float32)
    vols = [rng.standard_normal((n, n, n)).astype(np.float32) for _ in range(P)]
    mask = rng.uniform(0.0, 1.0, (n, n, n)).astype(np.float32)
    out = []
    for parts, m in ((vols, mask), ([v * mask for v in vols], None), (vols, np.ones((n, n, n), np.float32)),
                     (vols, None)):
        al = BatchAligner(n, np.zeros((n, n, n), np.float32), bp)
        al.set_particles(parts, [None] * P, [np.eye(3)] * P, mask=m)
        out.append((al.Q.get(), al.N.get(), al.A.get()))
    for a, b in ((out[0], out[1]), (out[2], out[3])):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_preparation_of_stacked_particles_reads_them_in_place(ctx):
    """Unmasked particles stacked at a stride of one volume go to the R2C
    without a staging copy (the ifftshift as the spectrum's sign): Q, N and
    A bit-identical to the same particles at a ragged stride (staged, 19 =
    16 + 2 + 1 chunks), the stack left unchanged, and Q against the FP64
    oracle's preparation; odd n (ifftshift staged) against the oracle."""
    from emclarity_amd.device import DeviceArray
    for n in (32, 15):
        P = 19
        rng = np.random.default_rng(12 + n)
        bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5)
        vols = rng.standard_normal((P, n, n, n)).astype(np.float32)
        nb = 4 * n ** 3
        stack = DeviceArray.from_numpy(vols, ctx=ctx)
        gh = np.zeros(P * (n ** 3 + 64), np.float32)
        for i in range(P):
            gh[i * (n ** 3 + 64): i * (n ** 3 + 64) + n ** 3] = vols[i].ravel()
        gap = DeviceArray.from_numpy(gh, ctx=ctx)
        out = []
        for views in ([stack.view((n, n, n), np.float32, nb * i) for i in range(P)],
                      [gap.view((n, n, n), np.float32, i * (nb + 256)) for i in range(P)]):
            al = BatchAligner(n, np.zeros((n, n, n), np.float32), bp.astype(np.float32))
            al.set_particles(views, [None] * P, [np.eye(3)] * P)
            out.append((al.Q.get(), al.N.get(), al.A.get()))
        for x, y in zip(*out):
            assert np.array_equal(x, y)
        assert np.array_equal(stack.get(), vols)
        Q = out[0][0]
        for i in (0, P - 1):
            ref = O.batch_prepare(vols[i].astype(np.float64), bp, None, al.support)
            # Q = a w (the oracle's weighted spectrum times its weight), in the
            # device layout [kz][kx][ky] of the half grid
            want = np.transpose((ref.a * ref.w)[:, :, : n // 2 + 1], (0, 2, 1)).reshape(-1)
            assert np.max(np.abs(Q[i] - want)) < 1e-5 * np.max(np.abs(want)), (n, i)
            assert abs(out[0][2][i] - ref.A) <= 1e-5 * ref.A


def test_preparation_in_64_particle_chunks(ctx):
    """70 stacked 64^3 particles with their own wedges prepare in chunks of
    64 + 4 + 2 (the R2C reading the stack in place): Q, N and A of the first
    and last particle of each chunk against the single-particle entry (Q to
    1e-5 of its maximum, N exact, A to 1e-6), A against the FP64 oracle."""
    import ctypes as C
    from emclarity_amd import _lib
    from emclarity_amd.device import DeviceArray
    n, P = 64, 70
    rng = np.random.default_rng(21)
    bp = O.bandpass3d((n, n, n), 1e-3, 40.0, 8.0, 2.5)
    vols = rng.standard_normal((P, n, n, n)).astype(np.float32)
    w = [BH_weightMaskMex((n, n, n), 1, S.tilt_table(55.0 + (i % 3) * 5.0, 3.0, 2.0 + 0.1 * i, 2.5),
                          [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]]).get() for i in range(4)]
    wedges = [w[i % 4] for i in range(P)]
    stack = DeviceArray.from_numpy(vols, ctx=ctx)
    views = [stack.view((n, n, n), np.float32, 4 * n ** 3 * i) for i in range(P)]
    al = BatchAligner(n, np.zeros((n, n, n), np.float32), bp.astype(np.float32))
    al.set_particles(views, wedges, [np.eye(3)] * P)
    Qb, Nb, Ab = al.Q.get(), al.N.get(), al.A.get()
    d_bp = DeviceArray.from_numpy(bp.astype(np.float32))
    for i in (0, 63, 64, 67, 68, 69):
        Q1 = DeviceArray((al.nh,), np.complex64)
        N1 = DeviceArray((al.nh,), np.float32)
        A1 = DeviceArray((1,), np.float64)
        wd = DeviceArray.from_numpy(wedges[i].astype(np.float32))
        _lib.call("emc_batch_prepare_particle", ctx.handle, n, views[i].vp, d_bp.vp, wd.vp, Q1.vp, N1.vp, A1.vp)
        q1, n1, a1 = Q1.get(), N1.get(), A1.get()[0]
        assert np.max(np.abs(Qb[i] - q1)) <= 1e-5 * np.max(np.abs(q1)), i
        assert np.array_equal(Nb[i], n1), i
        assert abs(Ab[i] - a1) <= 1e-6 * a1, i
        ref = O.batch_prepare(vols[i].astype(np.float64), bp, wedges[i], al.support)
        assert abs(Ab[i] - ref.A) <= 1e-5 * ref.A, i
    assert np.array_equal(stack.get(), vols)


def _host_refine(al, R0, s0, iters=12, step_angle=0.02, step_shift=0.3):
    """The refinement restated on the host (numpy FP64) around the device
    adjoint: the same batched BFGS emc_align_refine runs on the device."""
    P = al.n_part
    R = np.array(R0, dtype=np.float64).reshape(P, 3, 3)
    s = np.array(s0, dtype=np.float64).reshape(P, 3)
    f, g = al._eval_local(R, s)
    gw = np.linalg.norm(g[:, :3], axis=1) + 1e-12
    gs = np.linalg.norm(g[:, 3:], axis=1) + 1e-12
    H = np.zeros((P, 6, 6))
    H[:, [0, 1, 2], [0, 1, 2]] = (step_angle / gw)[:, None]
    H[:, [3, 4, 5], [3, 4, 5]] = (step_shift / gs)[:, None]
    H0 = H.copy()
    t = np.ones(P)
    for _ in range(iters):
        d = np.einsum("pij,pj->pi", H, g)
        bad = np.einsum("pi,pi->p", d, g) <= 0
        H[bad] = H0[bad]
        d[bad] = np.einsum("pij,pj->pi", H0[bad], g[bad])
        step = t[:, None] * d
        Rt = R @ rodrigues(step[:, :3])
        st = s + step[:, 3:]
        ft, gt = al._eval_local(Rt, st)
        ok = ft > f + 1e-4 * t * np.einsum("pi,pi->p", d, g)
        y = g - gt
        sy = np.einsum("pi,pi->p", step, y)
        for i in np.flatnonzero(ok & (sy > 1e-18)):
            V = np.eye(6) - np.outer(step[i], y[i]) / sy[i]
            H[i] = V @ H[i] @ V.T + np.outer(step[i], step[i]) / sy[i]
        R[ok], s[ok], f[ok], g[ok] = Rt[ok], st[ok], ft[ok], gt[ok]
        t[ok] = np.minimum(1.0, t[ok] * 2.0)
        t[~ok] *= 0.3
    return R, s, f


def test_device_refinement_matches_host_restatement(ctx):
    """emc_align_refine against the host BFGS driving the C-ABI adjoint:
    poses agree to 1e-6 and final CCs to 1e-7 except where an accept/reject
    decision falls on a rounding-level tie near the flat top, which sends
    the two trajectories apart (at most two particles of eight); every
    final CC agrees to 1e-6 (the same maximum).
    The device refinement forms the derivatives of its rotation increments
    from M (dM_a = G_a^T M, cross products in the brick-major kernel); the
    host passes the same derivatives as matrices, so the gradients agree to
    FP32 rounding, not bit for bit."""
    n = 48
    rng = np.random.default_rng(17)
    from emclarity_amd.workload import make_particle, make_size_class
    sc, blobs = make_size_class(n)
    parts = [make_particle(rng, sc, blobs) for _ in range(8)]
    wedges = [BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]]).get()
              for p in parts]
    al = BatchAligner(n, sc.ref, sc.bandpass, peak_radius=6)
    al.set_particles([p.data for p in parts], wedges, [p.start for p in parts])
    grid = eulerSearch("C1", 15, 6, 15, 6)
    score, shift = al.search(grid)
    best = np.argmax(score, axis=1)
    R0 = np.stack([parts[i].start @ local_matrix(*grid[best[i]]) for i in range(8)])
    s0 = shift[np.arange(8), best].astype(np.float64)
    Rd, sd, cd = al.refine(R0, s0, iters=10)
    Rh, sh, ch = _host_refine(al, R0, s0, iters=10)
    same = [np.max(np.abs(Rd[i] - Rh[i])) < 1e-6 and np.max(np.abs(sd[i] - sh[i])) < 1e-6 for i in range(8)]
    assert sum(same) >= 6, (same, cd, ch)
    for i in range(8):
        assert abs(cd[i] - ch[i]) < (1e-7 if same[i] else 1e-6) * abs(ch[i]), (i, cd[i], ch[i])
    f0, _ = al._eval_local(R0, s0)
    assert np.all(cd >= f0)   # only ascending steps are taken (deterministic adjoint)


def test_pick_matches_host(ctx):
    """emc_align_pick: first maximum (numpy argmax, ties included) and
    R0 = A @ defineMatrix([phi, theta, psi - phi], 'Bah', 'inv') in FP64."""
    import ctypes as C
    from emclarity_amd import _lib
    from emclarity_amd.device import DeviceArray
    rng = np.random.default_rng(11)
    P, NO = 7, 301
    score = rng.standard_normal((P, NO)).astype(np.float32)
    score[2, [5, 40]] = 9.0                       # a tie: the first wins
    shift = rng.uniform(-3, 3, (P, NO, 3)).astype(np.float32)
    starts = np.stack([O.define_matrix(rng.uniform(0, 360, 3), "Bah", "inv") for _ in range(P)])
    pts = np.stack([rng.uniform(0, 360, NO), rng.uniform(0, 15, NO), rng.uniform(-15, 15, NO)], 1)
    d_start = DeviceArray.from_numpy(np.ascontiguousarray(starts.transpose(0, 2, 1).reshape(P, 9)))
    d_pts = DeviceArray.from_numpy(np.ascontiguousarray(pts))
    R0 = DeviceArray((P, 9), np.float64)
    s0 = DeviceArray((P, 3), np.float64)
    best = DeviceArray((P,), np.int32)
    bs = DeviceArray((P,), np.float32)
    _lib.call("emc_align_pick", ctx.handle, P, NO, DeviceArray.from_numpy(score).vp, DeviceArray.from_numpy(shift).vp,
              d_start.vp, d_pts.vp, R0.vp, s0.vp, best.vp, bs.vp)
    b = np.argmax(score, axis=1)
    assert np.array_equal(best.get(), b) and b[2] == 5
    assert np.array_equal(bs.get(), score[np.arange(P), b])
    assert np.array_equal(s0.get(), shift[np.arange(P), b].astype(np.float64))
    for i in range(P):
        ph, th, ps = pts[b[i]]
        ref = starts[i] @ O.define_matrix([ph, th, ps - ph], "Bah", "inv")
        assert np.max(np.abs(R0.get()[i].reshape(3, 3) - ref)) < 1e-12


def test_refinement_graph_replays_match_eager():
    """emc_align_refine replays its iteration chain as a HIP graph: every
    replay (the same buffers: the cached graph; new buffers: a new capture)
    with eager work and host waits in between equals the eager launches
    exactly (the same kernels in the same order, deterministic sums)."""
    from emclarity_amd import autotune
    from emclarity_amd.device import Context, DeviceArray
    from emclarity_amd.workload import make_particle, make_size_class
    c = Context(0)
    n = 32
    rng = np.random.default_rng(23)
    sc, blobs = make_size_class(n)
    parts = [make_particle(rng, sc, blobs) for _ in range(5)]
    wedges = [BH_weightMaskMex((n, n, n), 1, p.tlt, [n // 2 + 1] * 3, [[n, n, n], [0, 0, 0]], ctx=c).get()
              for p in parts]
    al = BatchAligner(n, sc.ref, sc.bandpass, peak_radius=6, ctx=c)
    al.set_particles([p.data for p in parts], wedges, [p.start for p in parts])
    grid = eulerSearch("C1", 15, 6, 15, 6)
    score, shift = al.search(grid)
    best = np.argmax(score, axis=1)
    R0 = np.stack([parts[i].start @ local_matrix(*grid[best[i]]) for i in range(5)])
    s0 = shift[np.arange(5), best].astype(np.float64)
    autotune.apply(c, n, {"refine_graph": 0})
    Re, se, ce = al.refine(R0, s0, iters=8)
    autotune.apply(c, n, {"refine_graph": 1})
    for rep in range(3):                                  # new buffers each call
        Rg, sg, cg = al.refine(R0, s0, iters=8)
        assert np.array_equal(Rg, Re) and np.array_equal(sg, se) and np.array_equal(cg, ce), rep
        al.search(grid[:20])                              # eager work between replays
        c.synchronize()
    R = DeviceArray((5, 9), np.float64, c)
    s = DeviceArray((5, 3), np.float64, c)
    import ctypes as C
    from emclarity_amd import _lib
    for rep in range(3):                                  # the same buffers: the cached graph
        for d, h in ((R, np.ascontiguousarray(R0.reshape(5, 9))), (s, np.ascontiguousarray(s0))):
            _lib.call("emc_memcpy_h2d", c.handle, d.vp, h.ctypes.data_as(C.c_void_p), C.c_size_t(h.nbytes))
        cc = al.refine_device(R, s, iters=8)
        assert np.array_equal(R.get().reshape(5, 3, 3), Re) and np.array_equal(s.get(), se), rep
        assert np.array_equal(cc.get(), ce), rep
        al._eval_local(R0, s0)                            # eager adjoint between replays
    c.close()


def test_batched_entry_points_reject_bad_arguments(ctx):
    """Error paths of the batched and adjoint entry points: a box size
    outside 8..128, an empty batch, a support radius outside (0, n/2), a
    peak radius other than 4/6/8, an unknown table format, null pointers and
    refinement steps <= 0 each raise EmcError before any launch; a good call
    follows on the same context."""
    from emclarity_amd import _lib
    from emclarity_amd.device import DeviceArray
    buf = DeviceArray((4096,), np.float64)      # any non-null device pointer: no call reaches a launch
    b = buf.vp
    pts = _lib.d_arr([0.0, 0.0, 0.0])
    bad_batched = [  # n, n_part, table, n_orient, peak_radius, support_radius
        (6, 2, 0, 1, 6, 2.0), (300, 2, 0, 1, 6, 2.0), (16, 0, 0, 1, 6, 2.0), (16, 2, 0, 0, 6, 2.0),
        (16, 2, 0, 1, 5, 2.0), (16, 2, 0, 1, 6, 8.0), (16, 2, 0, 1, 6, 0.0), (16, 2, 7, 1, 6, 2.0)]
    for n, npart, tab, no, pr, sr in bad_batched:
        with pytest.raises(_lib.EmcError):
            _lib.call("emc_align_batched", ctx.handle, n, npart, tab, b, b, b, b, b, pts, no, pr, sr, b, b)
    with pytest.raises(_lib.EmcError):
        _lib.call("emc_align_batched", ctx.handle, 16, 2, 0, b, None, b, b, b, pts, 1, 6, 2.0, b, b)
    for n, npart, tab, sr in ((6, 2, 0, 2.0), (16, 0, 0, 2.0), (16, 2, 3, 2.0), (16, 2, 0, 8.5)):
        with pytest.raises(_lib.EmcError):
            _lib.call("emc_align_eval_grad", ctx.handle, n, npart, tab, b, b, b, b, b, b, sr, b)
    with pytest.raises(_lib.EmcError):
        _lib.call("emc_align_eval_grad", ctx.handle, 16, 2, 0, b, b, b, b, None, b, 2.0, b)
    for iters, sa, ss, sr in ((-1, 0.01, 0.1, 2.0), (4, 0.0, 0.1, 2.0), (4, 0.01, -0.1, 2.0), (4, 0.01, 0.1, 9.0)):
        with pytest.raises(_lib.EmcError):
            _lib.call("emc_align_refine", ctx.handle, 16, 2, 0, b, b, b, b, b, b, iters, sa, ss, sr, b)
    for n, tab in ((6, 0), (16, 9)):
        with pytest.raises(_lib.EmcError):
            _lib.call("emc_batch_reference_spectrum", ctx.handle, n, b, tab, b)
    with pytest.raises(_lib.EmcError):
        _lib.call("emc_align_pick", ctx.handle, 0, 1, b, b, b, b, b, b, None, None)
    # a good call: one particle, one orientation (0, 0, 0) from the identity start
    score = DeviceArray.from_numpy(np.array([0.5], np.float32))
    shift = DeviceArray.from_numpy(np.array([1.0, -2.0, 0.5], np.float32))
    start = DeviceArray.from_numpy(np.eye(3).ravel())
    apts = DeviceArray.from_numpy(np.zeros(3))
    R0 = DeviceArray((9,), np.float64)
    s0 = DeviceArray((3,), np.float64)
    _lib.call("emc_align_pick", ctx.handle, 1, 1, score.vp, shift.vp, start.vp, apts.vp, R0.vp, s0.vp, None, None)
    assert np.max(np.abs(R0.get() - np.eye(3).ravel())) < 1e-12
    assert np.array_equal(s0.get(), [1.0, -2.0, 0.5])
