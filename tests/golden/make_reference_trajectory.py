"""Writes reference_shader_trajectory.npz: gFusion's frame loop with the
reference's own compute shaders on Mesa's llvmpipe (oracle.reference.
RefSession; its host loop restated by the C oracle) over the first 16 frames
of synthetic session 0, 128^3 RG16I over 2 m, truncation 3 voxels, ICP
10/5/4, Track with recoverPose on (kinect.cpp's defaults).

Stored per frame: the pose after the frame (camera -> world, row-major), the
tracking decision and the alignment energy of the last ICP iteration; after
the last frame, the number of voxels with weight > 0, the sum of their
weights and the raycast hit count.  These are the reference's outputs (test
data): tests/test_gpu_reference_trajectory.py compares the HIP frame loop
with them.  Needs oracle/_ref/libkfref.so (python oracle/build_ref.py).

    python tests/golden/make_reference_trajectory.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oglfusionkinectv2_amd import kinect_config, synth  # noqa: E402
from oracle import reference as R  # noqa: E402

VOLUME, FRAMES = 128, 16
OUT = os.path.join(HERE, "reference_shader_trajectory.npz")


def main():
    depth, gt = synth.sequence(0, FRAMES)
    mu = kinect_config(VOLUME, 2.0, "f32").mu
    s = R.RefSession(VOLUME, 2.0, mu)
    s.reset(gt[0])
    poses = np.zeros((FRAMES, 4, 4), np.float32)
    tracked = np.zeros(FRAMES, np.int8)
    ae = np.zeros(FRAMES, np.float32)
    for f in range(FRAMES):
        tracked[f] = s.frame(depth[f].astype(np.float32))
        poses[f] = s.pose()
        ae[f] = s.ae
    vol = s.volume()
    w = vol[..., 1].astype(np.int64)
    hits = int((s.map("ref_vertex")[..., 3] > 0).sum())
    np.savez_compressed(OUT, poses=poses, tracked=tracked, ae=ae, gt=gt.astype(np.float32),
                        weighted_voxels=np.int64((w > 0).sum()), weight_sum=np.int64(w.sum()),
                        raycast_hits=np.int64(hits), renderer=np.array(R.RefShaders.info()))
    print("wrote", OUT, "tracked", tracked.tolist(), "final error vs gt %.4f m"
          % np.linalg.norm(poses[-1][:3, 3] - gt[FRAMES - 1][:3, 3]))


if __name__ == "__main__":
    main()
