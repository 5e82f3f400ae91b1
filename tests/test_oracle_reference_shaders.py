"""The C oracle against the reference's own compute shaders, executed by
Mesa's llvmpipe on the host (oracle/_ref/libkfref.so, built by
oracle/build_ref.py from the reference tree; skipped when it is not built).

Each stage gets identical inputs on both sides and is compared at the bar the
driver's arithmetic allows: the GL compiler is free to contract a*b+c into an
FMA and to evaluate normalize() as x * inversesqrt(dot), the oracle does
neither, so dot-product outputs may differ by an ulp; everything discrete
(result codes, invalid sentinels, hit masks, the int16 TSDF) and every stage
without such freedom is compared bit for bit.
"""
import numpy as np
import pytest

import oracle
from oracle import reference as R
from oglfusionkinectv2_amd import kinect_config, synth

pytestmark = pytest.mark.skipif(not R.built(), reason="reference shaders not built (oracle/build_ref.py)")


@pytest.fixture(scope="module", autouse=True)
def _context():
    """The llvmpipe context, created only when a test here runs."""
    if not R.available():
        pytest.skip("llvmpipe reference context does not come up")

CAM = np.array((365.5, 365.5, 255.5, 211.5), np.float32)
N = 128
DIM = (2.0, 2.0, 2.0)


@pytest.fixture(scope="module")
def o():
    return oracle.Oracle("f32")


@pytest.fixture(scope="module")
def seq():
    return synth.sequence(0, 8)


@pytest.fixture(scope="module")
def fused(o, seq):
    """The oracle's session after 6 frames at 128^3 RG16I (a populated
    volume, reference maps from its raycast) and frame 6 as the next input."""
    depth, gt = seq
    mu = kinect_config(N, 2.0, "f32").mu
    s = oracle.OracleSession(o, N, 2.0, mu, oracle.RG16I)
    s.set_mode(1, 0)
    s.reset(gt[0])
    for f in range(6):
        s.frame(depth[f].astype(np.float32))
    d = depth[6].astype(np.float32)
    s.depth(d)
    return s, d, mu


@pytest.fixture(scope="module")
def gl():
    return R.RefShaders(512, 424, (N, N, N))


def test_context_is_llvmpipe():
    assert "llvmpipe" in R.RefShaders.info()


def test_mip_is_the_box_mean_up_to_filter_rounding(o, seq, gl):
    """glGenerateMipmap of the R32F depth texture (gFusion.cpp:508): the
    driver's filter is the 2x2 box mean of the oracle (zeros included) up to
    the rounding of its bilinear weights: within 2.5e-4 of the block's range
    plus an ulp, and exact on a constant block."""
    depth, _ = seq
    for f in (0, 3):
        gl.depth_to_vertex(depth[f].astype(np.float32), CAM)
        src = depth[f].astype(np.float32)
        for level in (1, 2):
            g = gl.map("depth", level)
            m = o.mip_depth(src)
            h, w = m.shape
            blk = src.reshape(h, 2, w, 2)
            rng = blk.max(axis=(1, 3)) - blk.min(axis=(1, 3))
            diff = np.abs(g - m)
            assert np.all(diff <= 2.5e-4 * rng + np.spacing(np.abs(m))), (f, level)
            assert np.array_equal(g[rng == 0], m[rng == 0])
            if level == 1:
                # the mean has dropouts averaged in: a block with one zero reads 3/4 of the rest
                one_zero = (blk == 0).sum(axis=(1, 3)) == 1
                assert one_zero.sum() > 100
                rest = blk.sum(axis=(1, 3))[one_zero] / 3
                tol = 2.5e-4 * rng[one_zero] + np.spacing(rest)
                assert np.all(np.abs(g[one_zero] - 0.75 * rest) <= tol)
                assert np.all(np.abs(g[one_zero] - rest) > 0.2 * rest)
            src = g


def test_depth_to_vertex_is_bit_exact(o, seq, gl):
    """depthToVert.cs with invK of camPams / 2^l, from the same depth level."""
    depth, _ = seq
    gl.depth_to_vertex(depth[1].astype(np.float32), CAM)
    for level in range(3):
        d = gl.map("depth", level)
        assert np.array_equal(gl.map("vertex", level), o.depth_to_vertex(d, CAM / (1 << level))), level


def test_vertex_to_normal_within_an_ulp(o, fused, gl):
    """vertToNorm.cs from identical vertex maps: the invalid sentinel (2,0,0,0)
    at exactly the same pixels, unit normals within 2 ulp of 1.0."""
    s, d, _ = fused
    gl.depth_to_vertex(d, CAM)
    for level in range(3):
        gl.set_map("vertex", s.map(1, level), level)
    gl.vertex_to_normal()
    for level in range(3):
        g, r = gl.map("normal", level), s.map(2, level)
        inv_g, inv_r = g[..., 0] == 2, r[..., 0] == 2
        assert np.array_equal(inv_g, inv_r)
        assert np.array_equal(g[inv_g], r[inv_r])
        assert np.array_equal(np.isnan(g), np.isnan(r))
        ok = ~inv_g & ~np.isnan(g).any(-1)
        assert np.abs(g[ok] - r[ok]).max() <= 2.5e-7, level


def test_track_records(o, fused, gl):
    """track.cs from identical maps and matrices: the result code of every
    pixel and the Jacobian of every valid pixel bit-exact, the residual within
    4e-9 m (a 3-term dot product the compiler may contract)."""
    s, d, _ = fused
    refV, refN = s.map(3), s.map(4)
    gl.set_map("ref_vertex", refV)
    gl.set_map("ref_normal", refN)
    pose = s.pose()
    view = o.track_view(pose, CAM)
    for level in range(3):
        gl.set_map("vertex", s.map(1, level), level)
        gl.set_map("normal", s.map(2, level), level)
    for level in range(3):
        n = (512 >> level) * (424 >> level)
        gl.track(level, view, pose)
        gr, ge, gJ = gl.track_records(n)
        rr, re, rJ = o.track(s.map(1, level), s.map(2, level), refV, refN, pose, CAM)
        assert np.array_equal(gr, rr), level
        valid = rr == 1
        assert valid.sum() > 1000
        assert set(np.unique(rr)) <= {-5, -4, -2, -1, 1}
        assert np.array_equal(gJ[valid], rJ[valid])
        assert np.abs(ge[valid] - re[valid]).max() <= 4e-9


def test_reduce_is_bit_exact(o, fused, gl):
    """reduce.cs (8 x 112 serial FP32 sums) from identical records: the
    oracle's sum order is the shader's, so every partial is bit-exact; on the
    tracked records of all three levels and on random records with every
    result code."""
    s, _, _ = fused
    pose = s.pose()
    refV, refN = s.map(3), s.map(4)
    for level in range(3):
        rr, re, rJ = o.track(s.map(1, level), s.map(2, level), refV, refN, pose, CAM)
        gl.set_track_records(rr, re, rJ)
        assert np.array_equal(gl.reduce(level), o.reduce(rr, re, rJ, 512 >> level, 424 >> level)), level
    rng = np.random.default_rng(7)
    n = 512 * 424
    res = rng.choice(np.array([1, 1, 1, -1, -2, -4, -5], np.int32), n)
    err = rng.standard_normal(n).astype(np.float32) * 0.01
    J = rng.standard_normal((n, 6)).astype(np.float32)
    gl.set_track_records(res, err, J)
    got, want = gl.reduce(0), o.reduce(res, err, J, 512, 424)
    assert np.array_equal(got, want)
    assert np.all(want[:, 28] == [(424 + 7 - g) // 8 * 512 for g in range(8)])


def test_integrate_first_frame_is_bit_exact(o, seq, gl):
    """integrate.cs into the zeroed volume: bit-exact, including the voxels
    left of / above the optical axis where pow(x/z, 2) has a negative base
    (llvmpipe evaluates it as the square) and the int16 stores that truncate
    toward zero."""
    depth, gt = seq
    mu = kinect_config(N, 2.0, "f32").mu
    d = depth[0].astype(np.float32)
    gl.depth_to_vertex(d, CAM)
    zero = np.zeros((N, N, N, 2), np.int16)
    gl.set_volume(zero)
    gl.integrate(o.inverse4(gt[0]), o.camera_matrix(CAM), mu, DIM)
    got = gl.volume()
    want = o.integrate(zero.copy(), oracle.RG16I, DIM, d, gt[0], CAM, mu)
    assert (want[..., 1] > 0).sum() > 10000
    assert np.array_equal(got, want)


def test_integrate_into_a_fused_volume(o, fused, gl):
    """integrate.cs into the oracle's 6-frame volume: at most 1e-3 of the
    touched voxels may differ (a residual that lands on the reset threshold
    after an FMA-contracted product)."""
    s, d, mu = fused
    vol = s.volume()
    pose = s.pose()
    gl.depth_to_vertex(d, CAM)
    gl.set_volume(vol)
    gl.integrate(o.inverse4(pose), o.camera_matrix(CAM), mu, DIM)
    got = gl.volume()
    want = o.integrate(vol.copy(), oracle.RG16I, DIM, d, pose, CAM, mu)
    touched = (want != vol).any(-1).sum()
    assert touched > 10000
    assert (got != want).any(-1).sum() <= 1e-3 * touched


def test_raycast(o, fused, gl):
    """raycast.cs over the oracle's volume: the same hit mask; surface points
    within 1e-4 m for all but 1e-4 of the hits (an ulp in an interpolated
    TSDF value can move the switch from the large to the small step, and with
    it the crossing found, by a step or more: one ray in 200k here); normals
    (central differences of the trilinear TSDF) within 5e-4 for 99.9% and
    within 5e-3 for all but 1e-4 of the hits."""
    s, _, mu = fused
    vol = s.volume()
    pose = s.pose()
    step = float(np.float32(2.0) / np.float32(N))
    large = float(np.float32(0.75) * np.float32(mu))
    gl.set_volume(vol)
    gl.raycast(o.mul4(pose, o.inverse_camera_matrix(CAM)), 0.1, 3.0, step, large, DIM)
    gV, gN = gl.map("ref_vertex"), gl.map("ref_normal")
    rV, rN = o.raycast(vol, oracle.RG16I, DIM, pose, CAM, 0.1, 3.0, step, large)
    hit = rV[..., 3] > 0
    assert hit.sum() > 100000
    assert np.array_equal(gV[..., 3] > 0, hit)
    assert np.array_equal(gV[~hit], rV[~hit]) and np.array_equal(gN[~hit], rN[~hit])
    dv = np.abs(gV[hit] - rV[hit]).max(-1)
    dn = np.abs(gN[hit] - rN[hit]).max(-1)
    assert (dv > 1e-4).sum() <= 1e-4 * hit.sum()
    assert np.quantile(dv, 0.999) <= 1e-5
    assert (dn > 5e-3).sum() <= 1e-4 * hit.sum() and np.quantile(dn, 0.999) <= 5e-4


def test_frame_loop_follows_the_oracle(seq):
    """The reference shaders driven through gFusion's frame loop (pyramid,
    Track, integrate, raycast) against the oracle's frame loop, 8 frames at
    128^3 RG16I: the same tracking decision every frame and poses within
    1 mm / 1e-3 (ulp-level differences compound through ICP)."""
    depth, gt = seq
    mu = kinect_config(N, 2.0, "f32").mu
    o32 = oracle.Oracle("f32")
    s = oracle.OracleSession(o32, N, 2.0, mu, oracle.RG16I)
    s.set_mode(1, 0)
    r = R.RefSession(N, 2.0, mu)
    s.reset(gt[0])
    r.reset(gt[0])
    for f in range(8):
        d = depth[f].astype(np.float32)
        ts, tr = s.frame(d), r.frame(d)
        assert ts == tr, f
        ps, pr = s.pose(), r.pose()
        assert np.abs(ps[:3, 3] - pr[:3, 3]).max() <= 1e-3, f
        assert np.abs(ps[:3, :3] - pr[:3, :3]).max() <= 1e-3, f
    assert np.linalg.norm(r.pose()[:3, 3] - gt[7][:3, 3]) < 0.02


def test_track_sdf_records(o, fused, gl):
    """trackSDF.cs from identical vertex maps, volume and camToWorld, over a
    record buffer holding random values (pixels the shader leaves unwritten
    keep them, on both sides): the Huber weight h and the distance D of every
    record bit-exact; the Jacobian (central differences of interpolated TSDF
    values, which amplify an ulp of contraction) within 2e-6 of its row's
    largest entry (16 ulp of the element at worst, 7e-7 of the row here)."""
    s, d, _ = fused
    vol, pose = s.volume(), s.pose()
    gl.set_volume(vol)
    gl.depth_to_vertex(d, CAM)
    rng = np.random.default_rng(3)
    for level in range(3):
        gl.set_map("vertex", s.map(1, level), level)
    for level in range(3):
        n = (512 >> level) * (424 >> level)
        init = rng.standard_normal((512 * 424, 8)).astype(np.float32)
        gl.set_sdf_records(init)
        gl.track_sdf(level, pose, DIM)
        got = gl.sdf_records()[:n]
        want = o.track_sdf(vol, oracle.RG16I, DIM, s.map(1, level), pose, buf=init[:n])
        assert np.array_equal(got[:, :2], want[:, :2]), level
        assert (want[:, 0] != init[:n, 0]).sum() > 1000
        row = np.maximum(np.abs(want[:, 2:]).max(-1, keepdims=True), 1)
        assert np.all(np.abs(got[:, 2:] - want[:, 2:]) <= 2e-6 * row), level


def test_reduce_sdf_is_bit_exact(o, fused, gl):
    """reduceSDF.cs (D^2 summed twice, reduceSDF.cs:52/60) from identical
    records: every partial bit-exact on all three levels."""
    s, _, _ = fused
    vol, pose = s.volume(), s.pose()
    for level in range(3):
        w, h = 512 >> level, 424 >> level
        rec = o.track_sdf(vol, oracle.RG16I, DIM, s.map(1, level), pose)
        gl.set_sdf_records(rec)
        assert np.array_equal(gl.reduce_sdf(level), o.reduce_sdf(rec, w, h)), level


def test_intensity_projection(o, fused, gl):
    """mip.cs over the oracle's volume: within 1e-5 (sums along each ray)."""
    s, _, mu = fused
    vol, pose = s.volume(), s.pose()
    step = float(np.float32(2.0) / np.float32(N))
    large = float(np.float32(0.75) * np.float32(mu))
    gl.set_volume(vol)
    got = gl.intensity_projection(o.mul4(pose, o.inverse_camera_matrix(CAM)), 0.1, 3.0, step, large, DIM)
    want = o.intensity_projection(vol, oracle.RG16I, DIM, pose, CAM, 0.1, 3.0, large)
    assert (want != 0).any(-1).sum() > 100000
    assert np.abs(got - want).max() <= 1e-5


def test_frame_loop_track_sdf_follows_the_oracle(seq):
    """TrackSDF as the tracker (track mode 2), 6 frames at 128^3: the same
    decision every frame, poses within 1 mm, the cumulative twist within 1e-3."""
    depth, gt = seq
    mu = kinect_config(N, 2.0, "f32").mu
    s = oracle.OracleSession(oracle.Oracle("f32"), N, 2.0, mu, oracle.RG16I)
    r = R.RefSession(N, 2.0, mu)
    s.set_mode(2, 1)
    r.set_mode(2, 1)
    s.reset(gt[0])
    r.reset(gt[0])
    for f in range(6):
        d = depth[f].astype(np.float32)
        assert s.frame(d) == r.frame(d), f
        assert np.abs(s.pose() - r.pose()).max() <= 1e-3, f
    assert np.abs(s.cum_twist() - r.cum_twist).max() <= 1e-3
    assert np.abs(s.cum_twist()).max() > 1e-3


def test_recover_pose_follows_the_oracle(seq):
    """recoverPose: after 5 frames of Track (one pose kept in the library) the
    pose is moved 12 cm off; Track fails on both sides, both recover through
    the library with TrackSDF, to poses within 1 mm of each other."""
    depth, gt = seq
    mu = kinect_config(N, 2.0, "f32").mu
    s = oracle.OracleSession(oracle.Oracle("f32"), N, 2.0, mu, oracle.RG16I)
    r = R.RefSession(N, 2.0, mu)
    s.reset(gt[0])
    r.reset(gt[0])
    for f in range(5):
        d = depth[f].astype(np.float32)
        s.frame(d)
        r.frame(d)
    assert len(s.library()) == len(r.library) >= 1
    d = depth[5].astype(np.float32)
    s.depth(d)
    r.depth(d)
    P = s.pose().copy()
    P[:3, 3] += np.float32(0.12)
    s.set_pose(P)
    r.set_pose(P)
    assert s.track() == r.track() == 0
    assert s.recover_pose() == r.recover_pose() == 1
    assert np.abs(s.pose() - r.pose()).max() <= 1e-3
    assert np.linalg.norm(r.pose()[:3, 3] - gt[5][:3, 3]) < 0.08
