"""Parity of point-to-volume tracking (TrackSDF, trackSDF.cs + reduceSDF.cs),
recoverPose and the pose library on the GPU against the C oracle, through the
C ABI.

Tolerances: the first iteration of a TrackSDF call evaluates the same FP32
per-pixel expressions on both sides from the same pose, so its 8x32 partial
sums agree to 1e-4 relative (sum order differs: DPP tree + FP64 against the
reference's serial FP32); later iterations start from poses that differ by
that rounding, so they are compared by the pose tolerance of the loop:
<= 1e-4 rad and <= 0.1 mm per call, with the same tracked decision.
"""
import numpy as np
import pytest

import oracle
from conftest import rigid

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
from oglfusionkinectv2_amd import GFusion, kinect_config  # noqa: E402

STORAGE_ORACLE = {"rg16i": oracle.RG16I, "f32": oracle.F32X2, "f16": oracle.F16X2}


def make(volume=128, storage="rg16i", iterations=(4, 5, 10)):
    g = GFusion()
    cfg = kinect_config(volume, 2.0, storage, iterations=iterations)
    g.setConfig(cfg)
    g.allocateBuffers()
    g.setOption(K.OPT_FAST_MATH, 0)
    return g, cfg


def planes(vol, storage):
    if storage == "f16":
        return vol[..., 0].view(np.float16), vol[..., 1].view(np.float16)
    return vol[..., 0], vol[..., 1]


def pose_close(P, Q, tag):
    """<= 0.1 mm and <= 1e-4 rad.  The angle comes from the antisymmetric part
    of P Q^T (the trace form is ill-conditioned near zero and reads the float
    poses' departure from orthonormality, ~1e-7, as a rotation)."""
    assert np.abs(P[:3, 3] - Q[:3, 3]).max() < 1e-4, (tag, P[:3, 3], Q[:3, 3])
    dR = P[:3, :3].astype(np.float64) @ Q[:3, :3].T.astype(np.float64)
    A = (dR - dR.T) / 2
    ang = np.arcsin(min(1.0, np.linalg.norm([A[2, 1], A[0, 2], A[1, 0]])))
    assert ang < 1e-4, (tag, ang)


def fixture(o32, seq20, storage="rg16i", frames=4, volume=128):
    """Oracle volume after `frames` tracked frames; GPU and oracle sessions
    then hold the same volume, the next frame's pyramid and the same pose /
    m_pose_eig (ground truth of that frame, moved by a few mm)."""
    depth, gt = seq20
    g, cfg = make(volume, storage)
    s = oracle.OracleSession(o32, volume, 2.0, cfg.mu, STORAGE_ORACLE[storage])
    s.reset(gt[0])
    for f in range(frames):
        s.frame(depth[f].astype(np.float32))
    vol = s.volume()
    start = gt[frames].astype(np.float32).copy()
    start[:3, 3] += np.float32([0.004, -0.003, 0.002])
    g.Reset(start)
    s.reset(start)
    g.setVolume(*planes(vol, storage))
    s.set_volume(vol)
    g.setPoseEig(start)
    s.set_pose_eig(start)
    d = depth[frames].astype(np.float32)
    g.depthToVertex(d)
    s.depth(d)
    return g, s


@pytest.mark.parametrize("storage", ["rg16i", "f32", "f16"])
def test_track_sdf_matches_oracle(o32, seq20, storage):
    print("case: TrackSDF 128^3 %s" % storage, flush=True)
    g, s = fixture(o32, seq20, storage)
    t_o = s.track_sdf()
    t_g = g.TrackSDF()
    st = g.sdfState()
    olog = s.sdf_log()
    n_o = int(sum(1 for r in olog if r[2] != 0 or r[3] != 0))
    assert t_g == bool(t_o)
    assert st["iterations"] == n_o
    # iteration 0: the same FP32 records on both sides
    np.testing.assert_allclose(st["log"][0, 2], olog[0, 2], rtol=1e-4)
    np.testing.assert_allclose(st["log"][0, 4:], olog[0, 4:], rtol=1e-3, atol=1e-7)
    pose_close(g.getPose(), s.pose(), "pose")
    pose_close(st["pose_eig"], s.pose_eig(), "pose_eig")
    np.testing.assert_allclose(st["cum_twist"], s.cum_twist(), rtol=1e-3, atol=1e-6)


def test_track_sdf_partials_first_iteration(o32, seq20):
    """One TrackSDF iteration at level 0 only: the 8x32 partials of that
    iteration (reduceSDF.cs layout) match the oracle's to 1e-4 relative."""
    print("case: TrackSDF partials", flush=True)
    depth, gt = seq20
    g, cfg = make(iterations=(1, 0, 0))
    s = oracle.OracleSession(o32, 128, 2.0, cfg.mu, oracle.RG16I, iterations=(1, 0, 0))
    s.reset(gt[0])
    s.set_mode(1, 0)
    for f in range(4):
        s.frame(depth[f].astype(np.float32))
    vol = s.volume()
    g.Reset(gt[4])
    s.reset(gt[4])
    g.setVolume(*planes(vol, "rg16i"))
    s.set_volume(vol)
    d = depth[4].astype(np.float32)
    g.depthToVertex(d)
    s.depth(d)
    s.track_sdf()
    g.TrackSDF()
    gp, op = g.sdfState()["partials"], s.sdf_partials()
    scale = np.abs(op).max(axis=0) + 1e-30
    assert (np.abs(gp - op) / scale).max() < 1e-4
    np.testing.assert_array_equal(gp[:, 28], op[:, 28])


def test_track_sdf_second_call_reads_the_same_stale_records(o32, seq20):
    """The last column of each level is summed from records earlier dispatches
    left in the buffer; a second TrackSDF after the first reads those and
    still matches."""
    print("case: TrackSDF twice", flush=True)
    g, s = fixture(o32, seq20)
    assert g.TrackSDF() == bool(s.track_sdf())
    start = s.pose_eig()
    g.setPoseEig(start)
    g.setPose(s.pose())
    assert g.TrackSDF() == bool(s.track_sdf())
    pose_close(g.getPose(), s.pose(), "second call")
    np.testing.assert_allclose(g.sdfState()["log"][0, 2], s.sdf_log()[0, 2], rtol=1e-4)


def test_pose_library_and_recover_pose(o32, seq20):
    """Frames with per-frame state sync build the same pose library on both
    sides; recoverPose then runs TrackSDF from the entries and ends in the
    same state."""
    print("case: pose library + recoverPose", flush=True)
    depth, gt = seq20
    g, cfg = make()
    s = oracle.OracleSession(o32, 128, 2.0, cfg.mu, oracle.RG16I)
    g.Reset(gt[0])
    s.reset(gt[0])
    for f in range(10):
        g.setVolume(*planes(s.volume(), "rg16i"))
        g.setReferenceMaps(s.map(3), s.map(4))
        g.setPose(s.pose())
        g.frame(depth[f].astype(np.float32))
        s.frame(depth[f].astype(np.float32))
    lib_g, lib_o = g.getPoseLibrary(), s.library()
    assert len(lib_g) == len(lib_o) >= 1
    np.testing.assert_allclose(lib_g, lib_o, atol=2e-4)
    g.setVolume(*planes(s.volume(), "rg16i"))
    g.setPose(s.pose())
    g.setPoseEig(s.pose_eig())
    found_o = s.recover_pose()
    found_g = g.recoverPose()
    assert found_g == bool(found_o)
    pose_close(g.getPose(), s.pose(), "recover")


def test_frame_recovers_after_a_lost_frame(o32, seq20):
    """An all-zero depth frame cannot be tracked (AE = 0): the frame loop then
    calls recoverPose, on the GPU inside the frame's launches."""
    print("case: frame loop lost frame -> recoverPose", flush=True)
    depth, gt = seq20
    g, cfg = make()
    s = oracle.OracleSession(o32, 128, 2.0, cfg.mu, oracle.RG16I)
    g.Reset(gt[0])
    s.reset(gt[0])
    for f in range(8):
        g.setVolume(*planes(s.volume(), "rg16i"))
        g.setReferenceMaps(s.map(3), s.map(4))
        g.setPose(s.pose())
        g.setPoseEig(s.pose_eig())
        d = depth[f].astype(np.float32)
        if f == 6:
            d = np.zeros_like(d)
        g.frame(d)
        s.frame(d)
        st, ost = g.status(), s.stats()
        assert st["tracked"] == ost["tracked"], f
        pose_close(g.getPose(), s.pose(), "frame %d" % f)
    assert len(s.library()) >= 1


@pytest.mark.parametrize("mode", [2, 3])
def test_frame_loop_track_modes(o32, seq20, mode):
    """kinect.cpp:666-674 with P2V (mode 2) or P2P then P2V (mode 3): per-frame
    parity from identical state."""
    print("case: frame loop track mode %d" % mode, flush=True)
    depth, gt = seq20
    g, cfg = make()
    g.setOption(K.OPT_TRACK_MODE, mode)
    s = oracle.OracleSession(o32, 128, 2.0, cfg.mu, oracle.RG16I)
    s.set_mode(mode, 1)
    g.Reset(gt[0])
    s.reset(gt[0])
    for f in range(8):
        g.setVolume(*planes(s.volume(), "rg16i"))
        g.setReferenceMaps(s.map(3), s.map(4))
        g.setPose(s.pose())
        g.setPoseEig(s.pose_eig())
        g.frame(depth[f].astype(np.float32))
        s.frame(depth[f].astype(np.float32))
        st, ost = g.status(), s.stats()
        assert st["tracked"] == ost["tracked"], f
        assert st["counter"] == ost["counter"] and st["reset"] == ost["reset"]
        pose_close(g.getPose(), s.pose(), "frame %d" % f)
        v, _ = g.getVolume()
        dv = np.abs(v.astype(np.int32) - s.volume()[..., 0].astype(np.int32)) / 32766.0
        assert (dv > 1e-3).mean() < 1e-3, f


def test_track_sdf_graph_frames_equal_direct(seq20):
    """Frames with TrackSDF and recovery replayed from the frame graph give
    the same bits as the same launches issued directly."""
    print("case: TrackSDF graph vs direct", flush=True)
    depth, gt = seq20
    out = []
    for graph in (1, 0):
        g, _ = make()
        g.setOption(K.OPT_TRACK_MODE, 3)
        g.setOption(K.OPT_GRAPH, graph)
        g.Reset(gt[0])
        for f in range(8):
            g.frame(depth[f].astype(np.float32))
        out.append((g.getPose(), g.getVolume()[0], g.getVerts(), g.getPoseLibrary()))
        g.close()
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)
