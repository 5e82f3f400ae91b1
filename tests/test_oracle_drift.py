"""The oracle trajectories of tests/golden/oracle_trajectories.npz (written by
tests/golden/make_trajectories.py): the file is the oracle's output (its
first frames are recomputed here bit for bit), and it records how the
reference's frame loop behaves on the synthetic sequences -- at 128^3 (mu =
47 mm) it stays within centimetres of ground truth, at 512^3 with the same
3-voxel truncation (mu = 11.7 mm) it loses the scene within tens of frames,
in FP64 as in FP32, so the loss is the reference's TSDF rule
(integrate.cs:89-99, values >= 0.2 reset to zero) at that truncation and not
the arithmetic."""
import os

import numpy as np
import pytest

import oracle
from oglfusionkinectv2_amd import kinect_config, synth

HERE = os.path.dirname(os.path.abspath(__file__))
NPZ = os.path.join(HERE, "golden", "oracle_trajectories.npz")

pytestmark = pytest.mark.skipif(not os.path.exists(NPZ), reason="oracle_trajectories.npz not generated")


@pytest.fixture(scope="module")
def traj():
    return dict(np.load(NPZ))


def _err(P, gt):
    return np.linalg.norm(P[:, :3, 3] - gt[: len(P), :3, 3], axis=1)


def test_fixture_is_the_oracles_output(traj):
    depth, gt = synth.sequence(0, 12)
    np.testing.assert_array_equal(gt.astype(np.float64), traj["gt"][:12])
    o = oracle.Oracle("f32")
    s = oracle.OracleSession(o, 128, 2.0, kinect_config(128, 2.0, "f32").mu, oracle.F32X2)
    s.reset(gt[0])
    for f in range(12):
        s.frame(depth[f].astype(np.float32))
        np.testing.assert_array_equal(s.pose(), traj["c1_128_f32_poses"][f].astype(np.float32), err_msg=str(f))
        assert s.stats()["tracked"] == traj["c1_128_f32_tracked"][f]


def test_128_tracks_the_ground_truth(traj):
    e = _err(traj["c1_128_f32_poses"], traj["gt"])
    assert e.max() < 0.03
    assert traj["c1_128_f32_tracked"][4:].mean() > 0.95


def test_512_with_three_voxel_truncation_drifts_in_fp32_and_fp64(traj):
    gt = traj["gt"]
    e32 = _err(traj["c2_512_f32_poses"], gt)
    e64 = _err(traj["c3_512_f64_poses"], gt)
    assert e32[49] > 0.05 and e64[49] > 0.05  # more than 5 cm off by frame 50
    assert e32[-1] > 0.5 and e64[-1] > 0.5
    # the two precisions follow each other far more closely than either
    # follows the ground truth at first
    d = np.linalg.norm(traj["c2_512_f32_poses"][:, :3, 3] - traj["c3_512_f64_poses"][:, :3, 3], axis=1)
    assert d[:20].max() < 0.25 * max(e64[:20].max(), 1e-3) + 0.002
