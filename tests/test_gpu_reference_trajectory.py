"""The HIP frame loop against the reference's own compute shaders: the
trajectory of gFusion's loop with depthToVert / vertToNorm / track / reduce /
integrate / raycast run by Mesa's llvmpipe (tests/golden/
reference_shader_trajectory.npz, 16 frames of synthetic session 0 at 128^3
RG16I, Track with recoverPose on).  Both the oracle-arithmetic mode (ICP sums
in reduce.cs's order, correctly rounded integrate) and the default kernels
must take the same tracking decision every frame and stay within 1 mm /
1e-3 of the reference's pose; the fused model's weighted-voxel count and the
last raycast's hit count within 1 %."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
from oglfusionkinectv2_amd import GFusion, kinect_config, synth  # noqa: E402

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_shader_trajectory.npz")


@pytest.mark.parametrize("oracle_arithmetic", [1, 0])
def test_frame_loop_follows_the_reference_shaders(oracle_arithmetic):
    fix = np.load(FIX)
    depth, gt = synth.sequence(0, 16)
    g = GFusion()
    g.setConfig(kinect_config(128, 2.0, "rg16i"))
    g.allocateBuffers()
    g.setOption(K.OPT_ICP_REFERENCE_ORDER, oracle_arithmetic)
    g.setOption(K.OPT_FAST_MATH, 1 - oracle_arithmetic)
    g.Reset(gt[0])
    try:
        for f in range(16):
            g.frame(depth[f].astype(np.float32))
            P = g.getPose()
            assert g.status()["tracked"] == fix["tracked"][f], f
            assert np.abs(P[:3, 3] - fix["poses"][f][:3, 3]).max() <= 1e-3, (f, P[:3, 3], fix["poses"][f][:3, 3])
            assert np.abs(P[:3, :3] - fix["poses"][f][:3, :3]).max() <= 1e-3, f
        _, w = g.getVolume()
        nw = int((w > 0).sum())
        assert abs(nw - int(fix["weighted_voxels"])) <= 1e-2 * int(fix["weighted_voxels"]), nw
        hits = int((g.getVerts()[..., 3] > 0).sum())
        assert abs(hits - int(fix["raycast_hits"])) <= 1e-2 * int(fix["raycast_hits"]), hits
    finally:
        g.close()
