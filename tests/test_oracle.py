"""Pins the CPU oracle (oracle/kf_oracle.c) against what the mathematics fixes:
closed forms, special cases that reduce to library routines (numpy.linalg,
numpy.float16, scipy Rotation), invariants and brute force on tiny inputs.
No GPU needed."""
import numpy as np
import pytest

import oracle
from conftest import rigid, rot, small_cam

CAM = (365.5, 365.5, 255.5, 211.5)


# ---------------------------------------------------------------- fp16 -----
def test_half_conversion_matches_numpy(o32):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.uniform(-0.2, 0.2, 4000), rng.uniform(-70000, 70000, 2000),
        np.float32(2.0) ** rng.integers(-30, 17, 500) * rng.choice([-1, 1], 500),
        [0.0, -0.0, 65504.0, 65519.9, 65520.0, 1e-8, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, np.inf, -np.inf],
    ]).astype(np.float32)
    # exact ties: halfway between two halves (rounds to even)
    h = np.arange(0, 0x7bff, 97, dtype=np.uint16).view(np.float16).astype(np.float32)
    nxt = (np.arange(0, 0x7bff, 97, dtype=np.uint16) + 1).view(np.float16).astype(np.float32)
    vals = np.concatenate([vals, (h + nxt) / 2])
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    got = np.array([o32.f2h(v) for v in vals], dtype=np.uint16)
    np.testing.assert_array_equal(got, want)
    back = np.array([o32.h2f(int(b)) for b in want], dtype=np.float32)
    np.testing.assert_array_equal(back, want.view(np.float16).astype(np.float32))


# ---------------------------------------------------------------- matrices -
def test_camera_matrices(o32):
    K = o32.camera_matrix(CAM)
    iK = o32.inverse_camera_matrix(CAM)
    assert K[0, 0] == 365.5 and K[0, 2] == 255.5 and K[1, 2] == 211.5 and K[3, 3] == 1
    np.testing.assert_allclose(K.astype(np.float64) @ iK.astype(np.float64), np.eye(4), atol=5e-5)


def test_inverse4_rigid_closed_form(o32, o64):
    rng = np.random.default_rng(1)
    for _ in range(20):
        M = rigid(*rng.uniform(-1, 1, 3), rng.uniform(-2, 2, 3)).astype(np.float64)
        Rt = M[:3, :3].T
        want = np.eye(4)
        want[:3, :3] = Rt
        want[:3, 3] = -Rt @ M[:3, 3]
        np.testing.assert_allclose(o32.inverse4(M), want, atol=2e-6)
        np.testing.assert_allclose(o64.inverse4(M) @ M, np.eye(4), atol=1e-13)
        G = rng.normal(size=(4, 4))
        np.testing.assert_allclose(o64.inverse4(G) @ G, np.eye(4), atol=1e-10)


def test_mul4_is_matrix_product(o32, o64):
    rng = np.random.default_rng(2)
    A, B = rng.normal(size=(4, 4)), rng.normal(size=(4, 4))
    np.testing.assert_allclose(o64.mul4(A, B), A @ B, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(o32.mul4(A, B), (A.astype(np.float32) @ B.astype(np.float32)), rtol=1e-5, atol=1e-5)


def test_pose_update_is_rx_ry_rz(o32, o64):
    """glm::eulerAngleXYZ = Rx(a) Ry(b) Rz(c) (intrinsic XYZ), translation x0..2,
    applied on the left of the pose (gFusion.cpp:700-705)."""
    rng = np.random.default_rng(3)
    for _ in range(10):
        x = rng.uniform(-0.3, 0.3, 6)
        P = rigid(*rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)).astype(np.float64)
        D = np.eye(4)
        D[:3, :3] = rot(x[3], x[4], x[5])
        D[:3, 3] = x[:3]
        np.testing.assert_allclose(o64.pose_update(x, P), D @ P, atol=1e-12)
        np.testing.assert_allclose(o32.pose_update(x, P), D @ P, atol=2e-6)


# ---------------------------------------------------------------- solve ----
def test_solve6_spd_matches_numpy(o64):
    rng = np.random.default_rng(4)
    for _ in range(20):
        J = rng.normal(size=(50, 6))
        C = J.T @ J
        b = rng.normal(size=6)
        np.testing.assert_allclose(o64.solve6(C, b), np.linalg.solve(C, b), rtol=1e-9, atol=1e-12)


def test_solve6_rank_deficient_is_min_norm(o64):
    """JacobiSVD::solve gives the minimum-norm least-squares solution."""
    rng = np.random.default_rng(5)
    for rank in (1, 3, 5):
        J = rng.normal(size=(rank, 6))
        C = J.T @ J
        b = rng.normal(size=6)
        want = np.linalg.lstsq(C, b, rcond=6 * np.finfo(float).eps)[0]
        np.testing.assert_allclose(o64.solve6(C, b), want, rtol=1e-6, atol=1e-9)
    np.testing.assert_array_equal(o64.solve6(np.zeros((6, 6)), np.ones(6)), np.zeros(6))


# ---------------------------------------------------------------- pyramid --
def test_mip_is_2x2_mean_including_zeros(o32):
    rng = np.random.default_rng(6)
    d = (rng.integers(0, 1000, (8, 12)) * 4).astype(np.float32)  # means exact
    m = o32.mip_depth(d)
    want = d.reshape(4, 2, 6, 2).mean(axis=(1, 3))
    np.testing.assert_array_equal(m, want)
    z = np.zeros((4, 4), np.float32)
    z[0, 0] = 1000
    assert o32.mip_depth(z)[0, 0] == 250  # invalid zeros are averaged in (driver mip)


def test_depth_to_vertex_closed_form(o32, o64):
    rng = np.random.default_rng(7)
    h, w = 24, 32
    cam = small_cam(w, h)
    d = rng.uniform(400, 4000, (h, w)).astype(np.float32)
    d[3, 4] = 0
    y, x = np.mgrid[0:h, 0:w]
    z = d.astype(np.float64) / 1000
    want = np.stack([z * (x - cam[2]) / cam[0], z * (y - cam[3]) / cam[1], z, 0 * z], -1)
    np.testing.assert_allclose(o32.depth_to_vertex(d, cam), want, rtol=2e-6, atol=1e-9)
    np.testing.assert_allclose(o64.depth_to_vertex(d, cam), want, rtol=1e-14, atol=1e-15)
    assert not o32.depth_to_vertex(d, cam)[3, 4].any()


def test_normals_of_a_plane(o64):
    """Any plane seen by the camera: every interior normal equals the plane
    normal oriented towards the camera; left/up borders are invalid."""
    h, w = 20, 28
    cam = small_cam(w, h)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    n = np.array([0.2, -0.3, -1.0])
    n /= np.linalg.norm(n)
    # plane n.p = -1.5 ; ray p = z * r  ->  z = -1.5 / (n.r)
    r = np.stack([(x - cam[2]) / cam[0], (y - cam[3]) / cam[1], np.ones_like(x)], -1)
    z = -1.5 / (r @ n)
    v = np.concatenate([z[..., None] * r, np.zeros((h, w, 1))], -1)
    N = o64.vertex_to_normal(v)
    assert (N[0, :, 0] == 2).all() and (N[:, 0, 0] == 2).all()
    np.testing.assert_allclose(N[1:, 1:, :3], np.broadcast_to(n, (h - 1, w - 1, 3)), atol=1e-9)
    assert (N[1:, 1:, 3] == 1).all()


def test_normal_quirks(o64):
    """down neighbour is read at (x-1, y+1) (vertToNorm.cs:27); any zero
    neighbour gives the sentinel (2,0,0,0); collinear neighbours give NaN."""
    h, w = 6, 6
    v = np.zeros((h, w, 4))
    y, x = np.mgrid[0:h, 0:w]
    v[..., 0], v[..., 1], v[..., 2] = x * 0.01, y * 0.01, 1.0
    v2 = v.copy()
    v2[3, 1, 2] = 0  # (x-1, y+1) of pixel (2, 2)
    N = o64.vertex_to_normal(v2)
    assert list(N[2, 2]) == [2, 0, 0, 0]
    v3 = v.copy()
    v3[3, 2, 2] = 0  # (x, y+1) of pixel (2, 2): not read
    assert o64.vertex_to_normal(v3)[2, 2, 3] == 1
    col = np.zeros((h, w, 4))
    col[..., 2] = 1.0  # all neighbours identical -> zero cross product
    assert np.isnan(o64.vertex_to_normal(col)[2, 2, :3]).all()


# ---------------------------------------------------------------- track ----
def _one_pixel_maps(v_cam, n_cam, ref_v, ref_n, w=8, h=8, px=(3, 3)):
    vert = np.zeros((h, w, 4))
    norm = np.zeros((h, w, 4))
    norm[..., 0] = 2  # all other pixels invalid
    vert[px[1], px[0], :3] = v_cam
    norm[px[1], px[0], :3] = n_cam
    norm[px[1], px[0], 3] = 1
    refV = np.zeros((h, w, 4))
    refN = np.zeros((h, w, 4))
    refV[..., :3] = ref_v
    refN[..., :3] = ref_n
    return vert, norm, refV, refN


def test_track_result_codes_and_jacobian(o64):
    cam = small_cam(8, 8)
    pose = rigid(0.01, -0.02, 0.03, [0.1, -0.05, 0.2]).astype(np.float64)
    # camera-space point on the optical axis of pixel (3,3)
    vc = np.array([(3 - cam[2]) / cam[0], (3 - cam[3]) / cam[1], 1.0]) * 1.2
    pw = pose[:3, :3] @ vc + pose[:3, 3]
    nw = pose[:3, :3] @ np.array([0, 0, -1.0])
    ref_v = pw + 0.01 * nw + np.array([0.002, -0.001, 0.0])
    maps = _one_pixel_maps(vc, [0, 0, -1.0], ref_v, nw)
    res, err, J = o64.track(*maps, pose, cam)
    i = 3 * 8 + 3
    assert res[i] == 1 and (res[np.arange(64) != i] == -1).all()
    assert err[i] == pytest.approx(nw @ (ref_v - pw), rel=1e-9)
    np.testing.assert_allclose(J[i], np.concatenate([nw, np.cross(pw, nw)]), rtol=1e-9, atol=1e-12)
    # -4: reference point farther than dist_threshold
    res, _, _ = o64.track(*_one_pixel_maps(vc, [0, 0, -1.0], pw + 0.2, nw), pose, cam)
    assert res[i] == -4
    # -5: normals disagree
    res, _, _ = o64.track(*_one_pixel_maps(vc, [0, 0, -1.0], ref_v, [1.0, 0, 0]), pose, cam)
    assert res[i] == -5
    # -3: reference normal carries the -2 marker
    res, _, _ = o64.track(*_one_pixel_maps(vc, [0, 0, -1.0], ref_v, [-2.0, 0, 0]), pose, cam)
    assert res[i] == -3
    # -2: projects outside the reference image
    far = _one_pixel_maps(np.array([5.0, 0, 1.0]), [0, 0, -1.0], ref_v, nw)
    res, _, _ = o64.track(*far, pose, cam)
    assert res[i] == -2


def test_track_self_consistency(o32):
    """Reference maps = the input maps moved to world by the exact pose:
    every valid association has zero error."""
    rng = np.random.default_rng(8)
    h, w = 24, 32
    cam = small_cam(w, h)
    d = rng.uniform(1400, 1600, (h, w)).astype(np.float32)
    v = o32.depth_to_vertex(d, cam)
    n = o32.vertex_to_normal(v)
    pose = np.eye(4, dtype=np.float32)
    refV, refN = v.copy(), n.copy()
    refN[n[..., 0] == 2] = 0
    res, err, J = o32.track(v, n, refV, refN, pose, cam)
    ok = res == 1
    assert ok.mean() > 0.5
    assert np.abs(err[ok]).max() < 1e-6


# ---------------------------------------------------------------- reduce ---
def test_reduce_brute_force_and_counts(o32, o64):
    rng = np.random.default_rng(9)
    w, h = 64, 40
    res = rng.choice([1, -1, -2, -4, -5], size=w * h, p=[0.5, 0.2, 0.1, 0.1, 0.1]).astype(np.int32)
    err = rng.normal(size=w * h) * 0.01
    J = rng.normal(size=(w * h, 6))
    parts = o64.reduce(res, err, J, w, h)
    yy = np.arange(w * h) // w
    iu = np.triu_indices(6)
    for g in range(8):
        sel = (yy % 8 == g)
        ok = sel & (res == 1)
        e, Jg = err[ok], J[ok]
        want = np.concatenate([[np.sum(e * e)], (e[:, None] * Jg).sum(0), np.einsum("ni,nj->ij", Jg, Jg)[iu]])
        np.testing.assert_allclose(parts[g, :28], want, rtol=1e-10, atol=1e-12)
        assert parts[g, 28] == sel.sum()  # every pixel of the group's rows counts once
        assert not parts[g, 29:].any()
    p32 = o32.reduce(res, err, J, w, h)
    np.testing.assert_allclose(p32, parts, rtol=1e-4, atol=1e-4)


def test_reduce_single_valid_pixel_and_all_invalid(o64):
    w, h = 16, 16
    res = np.full(w * h, -4, np.int32)
    parts = o64.reduce(res, np.zeros(w * h), np.zeros((w * h, 6)), w, h)
    assert not parts[:, :28].any() and (parts[:, 28] == 2 * w).all()
    b, C, ae = o64.get_reduction(parts)
    assert ae == 0 and not b.any() and not C.any()
    jv = np.array([0.1, -0.2, 0.3, 0.4, -0.5, 0.6])
    res[7 * w + 3] = 1
    err = np.zeros(w * h)
    err[7 * w + 3] = 0.05
    J = np.zeros((w * h, 6))
    J[7 * w + 3] = jv
    b, C, ae = o64.get_reduction(o64.reduce(res, err, J, w, h))
    np.testing.assert_allclose(C, np.outer(jv, jv), atol=1e-15)
    np.testing.assert_allclose(b, 0.05 * jv, atol=1e-15)
    assert ae == pytest.approx(np.sqrt(0.05 ** 2 / (2 * w)))


def test_get_reduction_rules(o64):
    rng = np.random.default_rng(10)
    parts = rng.normal(size=(8, 32))
    parts[:, 0] = np.abs(parts[:, 0])
    parts[:, 28] = rng.integers(1, 100, 8)
    b, C, ae = o64.get_reduction(parts, oracle.LAST_ROW)  # only row 7 survives
    np.testing.assert_array_equal(b, parts[7, 1:7])
    assert np.allclose(C, C.T) and C[0, 0] == parts[7, 7] and C[5, 5] == parts[7, 27]
    bs, Cs, aes = o64.get_reduction(parts, oracle.SUM_ALL)
    np.testing.assert_allclose(bs, parts[:, 1:7].sum(0))
    assert aes == pytest.approx(np.sqrt(parts[:, 0].sum() / parts[:, 28].sum()))


# ---------------------------------------------------------------- integrate
def _closed_form_plane(N, dim, d_mm, mu, cam, w, h):
    """Per-voxel first-frame TSDF (RG16I) for a fronto-parallel wall at depth
    d_mm seen from the identity camera, in float64 (reference semantics)."""
    vox = dim / N
    c = (np.arange(N) + 0.5) * vox
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    val = np.zeros((N, N, N))
    wt = np.zeros((N, N, N))
    with np.errstate(divide="ignore", invalid="ignore"):
        px = cam[0] * X / Z + cam[2] + 0.5
        py = cam[1] * Y / Z + cam[3] + 0.5
    live = (Z >= 1e-4) & (px >= 0) & (px <= w - 1) & (py >= 0) & (py <= h - 1)
    diff = (d_mm / 1000 - Z) * np.sqrt(1 + (X / Z) ** 2 + (Y / Z) ** 2)
    sdf = np.minimum(1, diff / mu)
    upd = live & (diff > -mu) & (sdf < 0.2)
    val[upd] = np.clip(sdf[upd], -0.1, 0.2)
    wt[upd] = 1
    return np.trunc(val * 32766).astype(np.int16), wt.astype(np.int16), live, diff


def test_integrate_first_frame_closed_form(o32):
    N, dim, w, h = 24, 1.2, 64, 48
    cam = small_cam(w, h)
    mu = 0.1
    depth = np.full((h, w), 800.0, np.float32)
    vol = o32.new_volume((N, N, N), oracle.RG16I)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = (0.6, 0.6, 0.0)  # volume centred on the optical axis
    o32.integrate(vol, oracle.RG16I, (dim,) * 3, depth, pose, cam, mu)
    # same geometry in the camera frame: shift the grid
    vox = dim / N
    c = (np.arange(N) + 0.5) * vox
    Z, Y, X = np.meshgrid(c, c - 0.6, c - 0.6, indexing="ij")
    with np.errstate(divide="ignore", invalid="ignore"):
        px = cam[0] * X / Z + cam[2] + 0.5
        py = cam[1] * Y / Z + cam[3] + 0.5
    live = (Z >= 1e-4) & (px >= 0) & (px <= w - 1) & (py >= 0) & (py <= h - 1)
    diff = (0.8 - Z) * np.sqrt(1 + (X / Z) ** 2 + (Y / Z) ** 2)
    sdf = np.minimum(1, diff / mu)
    upd = live & (diff > -mu) & (sdf < 0.2)
    want_v = np.where(upd, np.trunc(np.clip(sdf, -0.1, 0.2) * 32766), 0)
    # voxels whose classification sits on a threshold are excluded (float32 vs float64)
    edge = (np.abs(diff + mu) < 1e-5) | (np.abs(sdf - 0.2) < 1e-5) | (np.abs(px - np.round(px)) < 1e-4)
    got_v = vol[..., 0].astype(np.float64)
    assert (np.abs(got_v - want_v)[~edge] <= 1).all()
    np.testing.assert_array_equal(vol[..., 1][~edge], upd[~edge].astype(np.int16))
    assert upd.sum() > 100 and (live & ~upd).sum() > 100  # both branches exercised


def test_integrate_weight_cap_and_zero_depth(o32):
    N, dim, w, h = 16, 1.0, 32, 24
    cam = small_cam(w, h)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = (0.5, 0.5, 0.0)
    vol = o32.new_volume((N, N, N), oracle.F32X2)
    depth = np.full((h, w), 700.0, np.float32)
    for _ in range(105):
        o32.integrate(vol, oracle.F32X2, (dim,) * 3, depth, pose, cam, 0.1, 100.0)
    assert vol[..., 1].max() == 100  # min(w + 1, maxWeight)
    before = vol.copy()
    o32.integrate(vol, oracle.F32X2, (dim,) * 3, np.zeros((h, w), np.float32), pose, cam, 0.1)
    np.testing.assert_array_equal(vol, before)  # depth 0 -> untouched
    pose2 = pose.copy()
    pose2[2, 3] = 5.0  # whole volume behind the camera
    o32.integrate(vol, oracle.F32X2, (dim,) * 3, depth, pose2, cam, 0.1)
    np.testing.assert_array_equal(vol, before)


def test_integrate_rg16i_truncates_toward_zero(o32):
    """set() stores int(v * 32766): -0.05 -> -1638 (not -1639)."""
    N, dim, w, h = 8, 0.8, 16, 16
    cam = small_cam(w, h)
    vol = o32.new_volume((N, N, N), oracle.RG16I)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = (0.4, 0.4, 0.0)
    o32.integrate(vol, oracle.RG16I, (dim,) * 3, np.full((h, w), 500.0, np.float32), pose, cam, 1.0)
    v = vol[..., 0][vol[..., 1] > 0].astype(np.int64)
    vf = o32.new_volume((N, N, N), oracle.F32X2)
    o32.integrate(vf, oracle.F32X2, (dim,) * 3, np.full((h, w), 500.0, np.float32), pose, cam, 1.0)
    f = vf[..., 0][vf[..., 1] > 0]
    np.testing.assert_array_equal(v, np.trunc(f.astype(np.float32) * np.float32(32766)).astype(np.int64))
    assert (f < 0).any() and (f > 0).any()


# ---------------------------------------------------------------- raycast --
def _plane_tsdf(N, dim, z0, mu, mode):
    c = (np.arange(N) + 0.5) * dim / N
    d = np.clip((z0 - c) / mu, -0.1, 0.2)
    vol = np.zeros((N, N, N, 2), np.float32)
    vol[..., 0] = d[:, None, None]
    vol[..., 1] = 1
    return vol


def test_raycast_analytic_plane(o32):
    N, dim = 48, 2.0
    mu = 3 * dim / N
    vol = _plane_tsdf(N, dim, 1.4, mu, oracle.F32X2)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = (1.0, 1.0, 0.0)
    rv, rn = o32.raycast(vol, oracle.F32X2, (dim,) * 3, pose, CAM, 0.1, 3.0, dim / N, 0.75 * mu)
    hit = rv[..., 3] > 0
    assert hit.mean() > 0.3
    assert np.abs(rv[hit][:, 2] - 1.4).max() < 0.5 * dim / N
    np.testing.assert_allclose(rn[hit][:, :3], np.broadcast_to([0, 0, -1], (hit.sum(), 3)), atol=1e-6)
    # hit points lie on their pixel's ray
    y, x = np.nonzero(hit)
    p = rv[hit][:, :3] - pose[:3, 3]
    np.testing.assert_allclose(p[:, 0] / p[:, 2], (x - CAM[2]) / CAM[0], atol=1e-5)


def test_raycast_miss_cases(o32):
    N, dim = 16, 1.0
    mu = 3 * dim / N
    vol = np.zeros((N, N, N, 2), np.float32)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = (3.0, 3.0, 3.0)  # box behind the camera
    rv, rn = o32.raycast(vol, oracle.F32X2, (dim,) * 3, pose, CAM, 0.1, 3.0, dim / N, 0.75 * mu, 64, 48)
    assert not rv.any() and not rn.any()
    vol[..., 0] = -0.1  # f(tnear) < 0: "already inside" -> no hit
    pose[:3, 3] = (0.5, 0.5, 0.0)
    rv, rn = o32.raycast(vol, oracle.F32X2, (dim,) * 3, pose, CAM, 0.1, 3.0, dim / N, 0.75 * mu, 64, 48)
    assert not rv.any()


def test_raycast_rg16i_equals_scaled_f32(o32):
    """RG16I interpolates the raw int16 and scales after (raycast.cs:76-81)."""
    N, dim = 32, 2.0
    mu = 3 * dim / N
    vf = _plane_tsdf(N, dim, 1.3, mu, oracle.F32X2)
    vi = np.zeros((N, N, N, 2), np.int16)
    vi[..., 0] = np.trunc(vf[..., 0] * 32766).astype(np.int16)
    vi[..., 1] = 1
    vf2 = vf.copy()
    vf2[..., 0] = vi[..., 0].astype(np.float32) / np.float32(32766)
    pose = np.eye(4, dtype=np.float32)
    pose[:3, 3] = (1.0, 1.0, 0.0)
    a = o32.raycast(vi, oracle.RG16I, (dim,) * 3, pose, CAM, 0.1, 3.0, dim / N, 0.75 * mu)
    b = o32.raycast(vf2, oracle.F32X2, (dim,) * 3, pose, CAM, 0.1, 3.0, dim / N, 0.75 * mu)
    hit = (a[0][..., 3] > 0) & (b[0][..., 3] > 0)
    assert hit.mean() > 0.3
    assert np.abs(a[0][hit] - b[0][hit]).max() < 1e-4


# ---------------------------------------------------------------- full loop
def test_frame_loop_tracks_and_gates(o32):
    from oglfusionkinectv2_amd import synth
    depth, gt = synth.sequence(0, 8)
    vox = np.float32(2.0) / np.float32(64)
    s = oracle.OracleSession(o32, 64, 2.0, float(vox * 3), oracle.RG16I)
    s.reset(gt[0])
    flags = []
    for f in range(8):
        flags.append(s.frame(depth[f].astype(np.float32)))
        st = s.stats()
        assert st["counter"] == f + 1
        assert st["reset"] == (1 if f < 3 else 0)
    assert flags[0] == 0 and all(flags[1:])  # empty model on frame 0 -> rejected
    assert np.linalg.norm(s.pose()[:3, 3] - gt[7][:3, 3]) < 0.05


def test_f32_and_f64_oracles_agree_on_first_frames(o32, o64):
    from oglfusionkinectv2_amd import synth
    depth, gt = synth.sequence(0, 4)
    vox = np.float32(2.0) / np.float32(64)
    a = oracle.OracleSession(o32, 64, 2.0, float(vox * 3), oracle.F32X2)
    b = oracle.OracleSession(o64, 64, 2.0, float(vox * 3), oracle.F64X2)
    a.reset(gt[0])
    b.reset(gt[0])
    for f in range(4):
        a.frame(depth[f].astype(np.float32))
        b.frame(depth[f].astype(np.float32))
    assert np.abs(a.pose()[:3, 3] - b.pose()[:3, 3]).max() < 1e-3
