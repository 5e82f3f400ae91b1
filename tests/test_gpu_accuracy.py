"""The 512^3 configurations over their 600 frames, free-running on the GPU,
against the oracle trajectories (tests/golden/oracle_trajectories.npz; FP16
TSDF with and without FP64 pose composition, FP32 TSDF): each GPU
trajectory stays as close to the FP64 oracle as the reference's own FP32
arithmetic does (the FP32 oracle against the FP64 oracle), within 2 mm at
frame 20, and its frames are accepted as often.  (The reference semantics
at 3-voxel truncation drift from ground truth in every precision; see
tests/test_oracle_drift.py.)"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
torch = pytest.importorskip("torch")

import bench  # noqa: E402


@pytest.mark.skipif(not os.path.exists(bench.ORACLE_TRAJ), reason="oracle trajectories not generated")
def test_600_frames_against_the_oracles():
    print("case: 512^3, 600 frames, f16 / f16 + FP64 pose / f32 against the oracle trajectories", flush=True)
    r = bench.accuracy_leg(torch, K.lib(), 512)
    print(json.dumps(r))
    own = r["oracle_f32_vs_f64"]["rms_mm"]
    ref = np.load(bench.ORACLE_TRAJ)
    oracle_tracked = float(ref["c3_512_f64_tracked"].mean())
    for leg in ("f16_fp64_pose", "f16", "f32"):
        a = r[leg]["vs_fp64_oracle"]
        assert a["rms_mm"] <= 1.5 * own + 5.0, (leg, a, own)
        assert a["frame_20_mm"] <= 2.0, (leg, a)
        assert abs(r[leg]["tracked_fraction"] - oracle_tracked) <= 0.01, leg
