"""Track() against the oracle at the benchmark volume sizes (64^3, 128^3,
512^3), from identical state: the oracle integrates the first frames, the GPU
session receives its reference maps (the raycast of that volume) and pose,
and both track the next frame.

- KF_OPT_ICP_REFERENCE_ORDER = 1 sums in reduce.cs's order
  (reduce.cs:27-127) and combines in FP32 (gFusion.cpp:1137-1167): the
  partials, AE, iteration count and pose are the oracle's bits.
- The default fused kernels sum the same products in another order (wave
  trees, FP64 group sums): pose within 0.1 mm / 1e-4 rad, AE within 1e-4
  relative.
"""
import numpy as np
import pytest

import oracle
from conftest import rigid

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
from oglfusionkinectv2_amd import GFusion, kinect_config  # noqa: E402

SIZES = [64, 128, 512]
_STATE = {}


def oracle_state(o32, seq20, volume, frames=3):
    """Oracle session (FP32 TSDF, last-row rule) after `frames` frames: its
    reference maps and pose, cached per volume size."""
    key = (volume, frames)
    if key not in _STATE:
        depth, gt = seq20
        cfg = kinect_config(volume, 2.0, "f32")
        s = oracle.OracleSession(o32, volume, 2.0, cfg.mu, oracle.F32X2)
        s.reset(gt[0])
        for f in range(frames):
            s.frame(depth[f].astype(np.float32))
        _STATE[key] = (s.map(3).copy(), s.map(4).copy(), s.pose().copy())
        del s
    return _STATE[key]


def track_pair(o32, seq20, volume, rule, ref_order, perturb, frames=3):
    """Track() of frame `frames` on the GPU and in the oracle from the same
    reference maps and starting pose."""
    depth, gt = seq20
    refV, refN, pose = oracle_state(o32, seq20, volume, frames)
    start = pose @ perturb
    cfg = kinect_config(volume, 2.0, "f32", reduce_rule=rule)
    g = GFusion()
    g.setConfig(cfg)
    g.allocateBuffers()
    g.setOption(K.OPT_ICP_REFERENCE_ORDER, ref_order)
    assert g.getOption(K.OPT_ICP_REFERENCE_ORDER) == ref_order
    g.Reset(np.eye(4, dtype=np.float32))
    g.setReferenceMaps(refV, refN)
    g.setPose(start)
    d = depth[frames].astype(np.float32)
    g.depthToVertex(d)
    # the oracle side: a fresh 8^3 session carries the track state only
    s = oracle.OracleSession(o32, 8, 2.0, cfg.mu, oracle.F32X2, rule=rule)
    s.reset(start)
    s.set_map(3, refV)
    s.set_map(4, refN)
    s.depth(d)
    t_gpu = g.Track()
    t_cpu = bool(s.track())
    out = dict(t_gpu=t_gpu, t_cpu=t_cpu, P=g.getPose(), Q=s.pose(), parts=g.readPartials(), oparts=s.partials(),
               ae=np.float32(g.alignmentEnergy()), oae=np.float32(s.stats()["ae"]),
               it=g.status()["iterations"], oit=s.stats()["iterations"])
    g.close()
    return out


PERTURB = {"none": np.eye(4, dtype=np.float32),
           "offset": rigid(0.004, -0.003, 0.002, [0.006, -0.004, 0.005])}


@pytest.mark.parametrize("volume", SIZES)
@pytest.mark.parametrize("rule", [oracle.LAST_ROW, oracle.SUM_ALL])
@pytest.mark.parametrize("start", ["none", "offset"])
def test_track_reference_order_is_bit_exact(o32, seq20, volume, rule, start):
    print("case: reference-order Track() %d^3 rule %d start %s" % (volume, rule, start), flush=True)
    r = track_pair(o32, seq20, volume, rule, 1, PERTURB[start])
    assert r["t_gpu"] == r["t_cpu"]
    assert r["it"] == r["oit"]
    np.testing.assert_array_equal(r["parts"], r["oparts"])
    assert r["ae"] == r["oae"]
    np.testing.assert_array_equal(r["P"], r["Q"])


@pytest.mark.parametrize("volume", SIZES)
@pytest.mark.parametrize("start", ["none", "offset"])
def test_track_default_order_within_tolerance(o32, seq20, volume, start):
    print("case: default Track() %d^3 start %s" % (volume, start), flush=True)
    r = track_pair(o32, seq20, volume, oracle.LAST_ROW, 0, PERTURB[start])
    assert r["t_gpu"] == r["t_cpu"]
    P, Q = r["P"], r["Q"]
    dt = np.abs(P[:3, 3] - Q[:3, 3]).max()
    dR = P[:3, :3] @ Q[:3, :3].T
    ang = np.arccos(np.clip((np.trace(dR) - 1) / 2, -1, 1))
    print("translation %.3g m, rotation %.3g rad" % (dt, ang))
    assert dt <= 1e-4 and ang <= 1e-4
    assert abs(float(r["ae"]) - float(r["oae"])) <= 1e-4 * max(float(r["oae"]), 1e-6)


@pytest.mark.parametrize("volume", [64, 128])
def test_frame_loop_reference_order(o32, seq20, volume):
    """The whole frame block with reference-order ICP sums and correctly
    rounded integrate: each frame's tracking decision, frame counter and pose
    against the oracle's (the raycast maps, which feed the next frame's ICP,
    may differ in a few pixels), the TSDF within 1e-3 of truncation outside
    <= 0.1 % of the voxels."""
    print("case: reference-order frame loop %d^3" % volume, flush=True)
    depth, gt = seq20
    cfg = kinect_config(volume, 2.0, "rg16i")
    g = GFusion()
    g.setConfig(cfg)
    g.allocateBuffers()
    g.setOption(K.OPT_FAST_MATH, 0)
    g.setOption(K.OPT_ICP_REFERENCE_ORDER, 1)
    g.Reset(gt[0])
    s = oracle.OracleSession(o32, volume, 2.0, cfg.mu, oracle.RG16I)
    s.reset(gt[0])
    worst = 0.0
    for f in range(8):
        g.frame(depth[f].astype(np.float32))
        s.frame(depth[f].astype(np.float32))
        st, ost = g.status(), s.stats()
        assert st["tracked"] == ost["tracked"] and st["counter"] == ost["counter"], f
        P, Q = g.getPose(), s.pose()
        d = np.abs(P[:3, 3] - Q[:3, 3]).max()
        worst = max(worst, d)
        assert d <= 1e-4, (f, d)
        g.setPose(Q)
    print("largest per-frame translation difference %.3g m" % worst)
    v, w = g.getVolume()
    ov = s.volume()
    bad = (np.abs(v.astype(np.int32) - ov[..., 0].astype(np.int32)) > 33) | (w != ov[..., 1])
    assert bad.mean() <= 1e-3
    g.close()


def test_reference_order_is_single_session_only():
    from oglfusionkinectv2_amd import batch
    g = GFusion()
    g.setConfig(kinect_config(64, 2.0, "f16"))
    g.allocateBuffers()
    g.setOption(K.OPT_ICP_REFERENCE_ORDER, 1)
    with pytest.raises(K.KfError):
        batch.SessionBatch([g])
    g.close()


def test_pose_fp64_composes_in_double(o32, o64, seq20):
    """KF_OPT_POSE_FP64: each ICP update is delta (double) times the FP64
    pose, then rounded: after two iterations the float pose is the FP64
    oracle's composition of the GPU's own two solutions (up to the last ulp of
    libm's cos / sin), the FP64 pose carried between them."""
    print("case: FP64 pose composition", flush=True)
    depth, gt = seq20
    refV, refN, pose = oracle_state(o32, seq20, 128, 3)
    start = (pose @ PERTURB["offset"]).astype(np.float32)
    g = GFusion()
    g.setConfig(kinect_config(128, 2.0, "f32"))
    g.allocateBuffers()
    g.setOption(K.OPT_POSE_FP64, 1)
    assert g.getOption(K.OPT_POSE_FP64) == 1
    g.Reset(np.eye(4, dtype=np.float32))
    g.setReferenceMaps(refV, refN)
    g.setPose(start)
    g.depthToVertex(depth[3].astype(np.float32))
    pd = start.astype(np.float64)
    for it in range(2):
        g.track(0, first=(it == 0), save_pose=(it == 0))
        x = g.readIcpSolution()
        pd = o64.pose_update(x, pd)
        got = g.getPose()
        np.testing.assert_allclose(got, pd.astype(np.float32), rtol=0, atol=2.5e-7)
    g.close()


def test_pose_fp64_precision_variant_holds_tolerances():
    """The precision variant (FP16 TSDF, FP64 pose composition) against the
    FP64 oracle: one frame from pinned state within the per-frame
    tolerances."""
    import json

    from precision import evaluate
    print("case: 128^3 f16 + FP64 pose against the FP64 oracle", flush=True)
    r = evaluate(volume=128, frames=9, storage="f16", pinned=(4, 8), against="f64",
                 with_f32_oracle=False, options={K.OPT_POSE_FP64: 1})
    print(json.dumps(r))
    for p in r["pinned"]:
        assert p["pose_mm"] <= 0.1 and p["pose_mrad"] <= 0.1, p
        assert p["tsdf"]["frac_gt_1e3"] <= 1e-3, p
