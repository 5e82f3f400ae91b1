"""Pins the oracle's point-to-volume tracking (TrackSDF, trackSDF.cs,
reduceSDF.cs, recoverPose, the pose library) against what the mathematics
fixes: exact trilinear reproduction of linear fields, the closed-form gradient
and Jacobian J = [p x g, g], float64 brute-force sums, numpy.linalg for the
LDL^T solve of SPD systems, the pseudo-inverse of the pivots for singular
diagonal systems, scipy.linalg.expm for the twist exponential.  No GPU."""
import numpy as np
import pytest
import scipy.linalg

import oracle

F32 = oracle.F32X2


def linear_volume(N, coef, dtype=np.float32):
    """(Z, Y, X, 2) volume with value a*i + b*j + c*k + d at voxel (i, j, k), weight 1."""
    a, b, c, d = coef
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    vol = np.zeros((N, N, N, 2), dtype)
    vol[..., 0] = a * i + b * j + c * k + d
    vol[..., 1] = 1
    return vol


# ---------------------------------------------------------------- SDF() ----
def test_sdf_sample_reproduces_linear_fields(o64):
    """Trilinear blending is exact for a linear field, also for the signed
    (extrapolating) fractions modf gives at negative coordinates."""
    rng = np.random.default_rng(10)
    coef = (0.01, -0.02, 0.005, 0.1)
    vol = linear_volume(16, coef, np.float64)
    for _ in range(200):
        loc = rng.uniform(-0.9, 14.0, 3)
        want = coef[0] * loc[0] + coef[1] * loc[1] + coef[2] * loc[2] + coef[3]
        got = o64.sdf_sample(vol, oracle.F64X2, loc)
        assert abs(got - want) < 1e-12, (loc, got, want)


def test_sdf_sample_uses_modf_not_floor(o64):
    """At x = -0.5 modf gives integer part 0 and fraction -0.5, so the corners
    are voxels 0 and 1 with weights 1.5 and -0.5 (floor would use -1 and 0)."""
    vol = np.zeros((4, 4, 4, 2))
    vol[0, 0, 0, 0] = 1.0
    vol[0, 0, 1, 0] = 3.0
    assert o64.sdf_sample(vol, oracle.F64X2, (-0.5, 0.0, 0.0)) == pytest.approx(1.5 * 1.0 - 0.5 * 3.0)
    assert o64.sdf_sample(vol, oracle.F64X2, (0.5, 0.0, 0.0)) == pytest.approx(2.0)
    # corners beyond the volume read zero
    assert o64.sdf_sample(vol, oracle.F64X2, (3.5, 0.0, 0.0)) == 0.0


def test_sdf_sample_rg16i_scales_each_corner(o32):
    vol = np.zeros((4, 4, 4, 2), np.int16)
    vol[1, 1, 1, 0] = 16383
    vol[1, 1, 2, 0] = -32766
    got = o32.sdf_sample(vol, oracle.RG16I, (1.25, 1.0, 1.0))
    s = np.float32(0.00003051944088)
    want = np.float32(16383) * s * np.float32(0.75) + np.float32(-32766) * s * np.float32(0.25)
    assert got == pytest.approx(float(want), rel=1e-6)


# ---------------------------------------------------------------- track ----
def test_track_sdf_pixel_closed_form(o64):
    """Linear TSDF f = a i + b j + c k + d (index space): D = f(cvp) with cvp =
    p * N / dim - (1, 1, 0.5) (trackSDF.cs:474), g = (a, b, c) / voxel,
    J = [p x g, g], huber weight 1 below |D| = 0.02 and 0.02/|D| above."""
    rng = np.random.default_rng(11)
    N, dim = 32, 2.0
    coef = (0.002, -0.003, 0.001, 0.0)
    vol = linear_volume(N, coef, np.float64)
    vox = dim / N
    for _ in range(50):
        T = np.eye(4)
        from scipy.spatial.transform import Rotation
        T[:3, :3] = Rotation.from_euler("XYZ", rng.uniform(-0.3, 0.3, 3)).as_matrix()
        T[:3, 3] = rng.uniform(0.6, 1.2, 3)
        v = rng.uniform(-0.2, 0.2, 3)
        p = T[:3, :3] @ v + T[:3, 3]
        cvp = p * N / dim - np.array([1.0, 1.0, 0.5])
        D = coef[0] * cvp[0] + coef[1] * cvp[1] + coef[2] * cvp[2] + coef[3]
        g = np.array(coef[:3]) / vox
        rec = o64.track_sdf_pixel(vol, oracle.F64X2, (dim,) * 3, T, v)
        h = 1.0 if abs(D) < 0.02 else 0.02 / abs(D)
        np.testing.assert_allclose(rec[:2], [h, D], rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(rec[2:5], np.cross(p, g), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(rec[5:8], g, rtol=1e-9, atol=1e-12)


def test_track_sdf_huber_and_far_values(o64):
    N, dim = 16, 2.0
    vol = linear_volume(N, (0.0, 0.0, 0.0, 0.5), np.float64)
    rec = o64.track_sdf_pixel(vol, oracle.F64X2, (dim,) * 3, np.eye(4) + np.diag([0, 0, 0, 0]), [0.9, 0.9, 0.9])
    assert rec[1] == pytest.approx(0.5) and rec[0] == pytest.approx(0.02 / 0.5)
    # |D| >= 1: the record is zero (trackSDF.cs:509-511, 546-556)
    vol2 = linear_volume(N, (0.0, 0.0, 0.0, 1.5), np.float64)
    rec2 = o64.track_sdf_pixel(vol2, oracle.F64X2, (dim,) * 3, np.eye(4), [0.9, 0.9, 0.9])
    assert np.all(rec2 == 0)


def test_track_sdf_skips_last_column(o32):
    rng = np.random.default_rng(12)
    N = 16
    vol = linear_volume(N, (0.001, 0.002, -0.001, 0.01))
    vert = np.zeros((5, 7, 4), np.float32)
    vert[..., :3] = rng.uniform(0.3, 1.5, (5, 7, 3))
    buf = np.full((5 * 7, 8), 7.0, np.float32)
    out = o32.track_sdf(vol, F32, (2.0,) * 3, vert, np.eye(4), buf)
    out = out.reshape(5, 7, 8)
    assert np.all(out[:, 6] == 7.0)
    assert not np.any(out[:, :6] == 7.0)


# ---------------------------------------------------------------- reduce ---
def test_reduce_sdf_brute_force(o32, o64):
    rng = np.random.default_rng(13)
    w, h = 40, 21
    buf = rng.normal(size=(h * w, 8))
    buf[:, 0] = rng.uniform(0, 1, h * w)
    out = o64.reduce_sdf(buf, w, h)
    for g in range(8):
        rows = buf.reshape(h, w, 8)[g::8].reshape(-1, 8)
        hh, D, J = rows[:, 0], rows[:, 1], rows[:, 2:]
        np.testing.assert_allclose(out[g, 0], 2 * np.sum(D * D), rtol=1e-12)
        np.testing.assert_allclose(out[g, 1:7], (hh[:, None] * J * D[:, None]).sum(0), rtol=1e-10, atol=1e-12)
        iu = np.triu_indices(6)
        JJ = (hh[:, None, None] * J[:, :, None] * J[:, None, :]).sum(0)
        np.testing.assert_allclose(out[g, 7:28], JJ[iu], rtol=1e-10, atol=1e-12)
        assert out[g, 28] == rows.shape[0]
        assert np.all(out[g, 29:] == 0)
    o = o32.reduce_sdf(buf.astype(np.float32), w, h)
    np.testing.assert_allclose(o, out, rtol=2e-4, atol=1e-4)


def test_reduce_sdf_single_pixel_is_outer_product(o64):
    w, h = 9, 8
    buf = np.zeros((h * w, 8))
    J = np.array([0.3, -0.2, 0.1, 1.0, 2.0, -0.5])
    buf[7 * w + 4] = [0.5, 0.2] + list(J)
    out = o64.reduce_sdf(buf, w, h)
    b, Cm, ae = o64.get_reduction(out, oracle.LAST_ROW)
    np.testing.assert_allclose(b, 0.5 * J * 0.2, rtol=1e-14)
    np.testing.assert_allclose(Cm, 0.5 * np.outer(J, J), rtol=1e-14)
    assert ae == pytest.approx(np.sqrt(2 * 0.04 / w))


# ---------------------------------------------------------------- solve ----
def test_ldlt_solve_spd_matches_numpy(o64):
    rng = np.random.default_rng(14)
    for _ in range(30):
        A = rng.normal(size=(6, 6))
        A = A @ A.T + rng.uniform(0.01, 1.0) * np.eye(6)
        b = rng.normal(size=6)
        np.testing.assert_allclose(o64.ldlt_solve6(A, b), np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


def test_ldlt_solve_pivots_and_pseudo_inverse(o64):
    """Zero pivots of a PSD diagonal system give zero components; the solve is
    permutation-invariant (pivoting), and all-zero matrices give zero."""
    A = np.diag([4.0, 0.0, 2.0, 0.0, 1.0, 0.0])
    b = np.array([4.0, 5.0, 2.0, 7.0, 1.0, 3.0])
    np.testing.assert_allclose(o64.ldlt_solve6(A, b), [1, 0, 1, 0, 1, 0], atol=1e-15)
    assert np.all(o64.ldlt_solve6(np.zeros((6, 6)), b) == 0)
    rng = np.random.default_rng(15)
    A = rng.normal(size=(6, 6))
    A = A @ A.T + np.eye(6)
    b = rng.normal(size=6)
    P = np.eye(6)[rng.permutation(6)]
    x = o64.ldlt_solve6(A, b)
    xp = o64.ldlt_solve6(P @ A @ P.T, P @ b)
    np.testing.assert_allclose(P.T @ xp, x, rtol=1e-10, atol=1e-12)
    # rank-deficient PSD with b in the range: A x = b holds
    U = rng.normal(size=(6, 3))
    A = U @ U.T
    b = A @ rng.normal(size=6)
    x = o64.ldlt_solve6(A, b)
    np.testing.assert_allclose(A @ x, b, rtol=1e-7, atol=1e-9)


def test_twist_exp_matches_expm(o32, o64):
    rng = np.random.default_rng(16)
    for scale in (1e-9, 1e-5, 1e-3, 0.1, 1.0, 2.5):
        for _ in range(5):
            xi = rng.normal(size=6) * scale
            M = np.zeros((4, 4))
            M[:3, :3] = [[0, -xi[2], xi[1]], [xi[2], 0, -xi[0]], [-xi[1], xi[0], 0]]
            M[:3, 3] = xi[3:]
            want = scipy.linalg.expm(M)
            np.testing.assert_allclose(o64.twist_exp(xi), want, rtol=1e-12, atol=1e-14)
            np.testing.assert_allclose(o32.twist_exp(xi), want, rtol=0, atol=2e-6 * max(1.0, scale))


def test_sdf_step_is_scaled_damped_solve(o64):
    rng = np.random.default_rng(17)
    for it in range(4):
        Jm = rng.normal(size=(50, 6))
        Cm = Jm.T @ Jm
        b = rng.normal(size=6)
        r0 = rng.normal(size=6) * 0.01
        cn, r, rp = o64.sdf_step(b, Cm, it, r0, r0 * 0.5)
        m = Cm.max()
        dx = np.linalg.solve(Cm / m + 0.01 * it * np.eye(6), b / m)
        np.testing.assert_allclose(r, r0 - dx, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(rp, r, rtol=0, atol=0)
        assert cn == pytest.approx(np.linalg.norm(r - r0 * 0.5), rel=1e-9)


# ---------------------------------------------------------------- session --
def _session(o, volume=64):
    vox = np.float32(2.0) / np.float32(volume)
    return oracle.OracleSession(o, volume, 2.0, float(vox * np.float32(3.0)), oracle.RG16I)


def test_reset_sets_pose_eig_to_translation_only(o32):
    from conftest import rigid
    s = _session(o32)
    P = rigid(0.1, -0.2, 0.3, [1.0, 1.1, 0.2])
    s.reset(P)
    want = np.eye(4, dtype=np.float32)
    want[:3, 3] = P[:3, 3]
    np.testing.assert_array_equal(s.pose_eig(), want)
    assert np.all(s.cum_twist() == 0) and len(s.library()) == 0


def test_pose_library_spacing_and_recover_with_empty_library(o32, seq20):
    depth, gt = seq20
    s = _session(o32)
    s.reset(gt[0])
    assert s.recover_pose() == 0  # empty library
    for f in range(6):
        s.frame(depth[f].astype(np.float32))
    lib = s.library()
    # frame 0 is never tracked (empty reference maps); the first success adds a pose
    assert len(lib) >= 1
    for i in range(len(lib)):
        for j in range(i):
            assert np.linalg.norm(lib[i][:3, 3] - lib[j][:3, 3]) >= 0.2
    np.testing.assert_array_equal(lib[0][3], [0, 0, 0, 1])


def test_track_sdf_converges_to_one_pose_from_different_starts(o32, seq20):
    """With the volume of a few tracked frames at 128^3, TrackSDF started from
    the frame's ground-truth pose moved by about 1 cm in different directions
    succeeds and lands on the same pose (within 0.5 mm), less than one voxel
    from the ground truth (the -1 / -1 / -0.5 voxel offset of trackSDF.cs:474
    against integrate.cs's voxel centres biases the optimum by a fraction of a
    voxel).  On success m_pose = m_pose_eig and m_cumTwist is the result."""
    depth, gt = seq20
    s = _session(o32, 128)
    vox = 2.0 / 128
    s.reset(gt[0])
    for f in range(4):
        s.frame(depth[f].astype(np.float32))
    s.depth(depth[4].astype(np.float32))
    ends = []
    for d in ([0.01, -0.006, 0.004], [-0.008, 0.004, -0.006]):
        start = gt[4].astype(np.float32).copy()
        start[:3, 3] += d
        s.set_pose_eig(start)
        twist0 = s.cum_twist()
        assert s.track_sdf() == 1
        P = s.pose()
        ends.append(P[:3, 3].copy())
        assert np.linalg.norm(P[:3, 3] - gt[4][:3, 3]) < vox
        np.testing.assert_array_equal(s.pose_eig()[:3], P[:3])
        log = s.sdf_log()
        n = sum(1 for r in log if r[2] != 0)
        np.testing.assert_allclose(s.cum_twist() - twist0, log[n - 1, 4:], rtol=1e-9, atol=1e-15)
    assert np.linalg.norm(ends[0] - ends[1]) < 5e-4


def test_sdf_step_on_a_zero_system_is_nan(o64):
    """C = 0: the scaling 1 / max(C) is infinite, the scaled system is NaN, and
    Eigen's LDLT (whole diagonal invalid: no factorisation, identity
    permutation) propagates NaN through the triangular solves -- so TrackSDF
    cannot converge on such an iteration (gFusion.cpp:890-931)."""
    with np.errstate(all="ignore"):
        cn, r, rp = o64.sdf_step(np.zeros(6), np.zeros((6, 6)), 0, np.zeros(6), np.zeros(6))
    assert np.isnan(cn) and np.isnan(r[:5]).all()


def test_track_sdf_on_an_empty_volume_fails_and_keeps_the_pose(o32, seq20):
    depth, gt = seq20
    s = _session(o32, 64)
    s.reset(gt[0])
    s.depth(depth[0].astype(np.float32))
    eig0 = s.pose_eig()
    assert s.track_sdf() == 0
    np.testing.assert_array_equal(s.pose(), gt[0].astype(np.float32))
    np.testing.assert_array_equal(s.pose_eig(), eig0)
    assert np.all(s.cum_twist() == 0)
