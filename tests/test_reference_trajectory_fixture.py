"""tests/golden/reference_shader_trajectory.npz (the frame loop of the
reference's own compute shaders on llvmpipe, 16 frames at 128^3 RG16I) pinned
on the CPU: the C oracle's frame loop follows it (same decisions, poses within
1 mm), and, where the reference library is built, its generator reproduces
its first frames bit for bit."""
import os

import numpy as np
import pytest

import oracle
from oracle import reference as R
from oglfusionkinectv2_amd import kinect_config, synth

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_shader_trajectory.npz")


@pytest.fixture(scope="module")
def fix():
    return np.load(FIX)


def test_fixture_shape(fix):
    assert fix["poses"].shape == (16, 4, 4)
    assert fix["tracked"][0] == 0 and fix["tracked"][1:].all()
    assert "llvmpipe" in str(fix["renderer"])
    assert np.linalg.norm(fix["poses"][-1][:3, 3] - fix["gt"][-1][:3, 3]) < 0.03


def test_oracle_follows_the_reference_shaders(fix):
    depth, gt = synth.sequence(0, 16)
    assert np.allclose(gt, fix["gt"], atol=1e-6)
    mu = kinect_config(128, 2.0, "f32").mu
    s = oracle.OracleSession(oracle.Oracle("f32"), 128, 2.0, mu, oracle.RG16I)
    s.reset(gt[0])
    for f in range(16):
        assert s.frame(depth[f].astype(np.float32)) == fix["tracked"][f], f
        assert np.abs(s.pose() - fix["poses"][f]).max() <= 1e-3, f
    w = s.volume()[..., 1].astype(np.int64)
    assert abs(int((w > 0).sum()) - int(fix["weighted_voxels"])) <= 1e-2 * int(fix["weighted_voxels"])


@pytest.mark.skipif(not R.built(), reason="reference shaders not built (oracle/build_ref.py)")
def test_generator_reproduces_the_fixture(fix):
    if not R.available():
        pytest.skip("llvmpipe reference context does not come up")
    depth, gt = synth.sequence(0, 4)
    r = R.RefSession(128, 2.0, kinect_config(128, 2.0, "f32").mu)
    r.reset(gt[0])
    for f in range(4):
        assert r.frame(depth[f].astype(np.float32)) == fix["tracked"][f]
        assert np.array_equal(r.pose(), fix["poses"][f]), f
