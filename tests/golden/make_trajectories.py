"""Writes oracle_trajectories.npz: the CPU oracle's free-running frame loop on
the seeded synthetic sequence of session 0, per configuration:

  c1_128_f32   128^3 FP32 TSDF (mu = 3 voxels), 100 frames  (FP32 oracle)
  c2_512_f32   512^3 FP32 TSDF (mu = 3 voxels), 600 frames  (FP32 oracle)
  c3_512_f64   512^3 FP64 TSDF (mu = 3 voxels), 600 frames  (FP64 oracle)

For each: the pose after every frame (camera -> world, row-major), the
tracking decision of every frame, and the ground-truth trajectory.  These are
the oracle's outputs (test data): the GPU path's trajectories are compared
with them (ATE against the FP64 oracle, the reference semantics' own drift
against ground truth, reported in oracle_drift.md), and
tests/test_oracle_drift.py re-runs the first frames to pin the file to the
oracle.

    python tests/golden/make_trajectories.py            # about 20 min on 8 cores
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oglfusionkinectv2_amd import kinect_config, synth  # noqa: E402

CONFIGS = {
    "c1_128_f32": (128, "f32", 100),
    "c2_512_f32": (512, "f32", 600),
    "c3_512_f64": (512, "f64", 600),
}
OUT = os.path.join(HERE, "oracle_trajectories.npz")


def run(volume, precision, frames, depth, gt, log=None):
    """(poses [F,4,4] float64, tracked [F] int8) of the oracle's frame loop."""
    o = oracle.Oracle(precision)
    mu = kinect_config(volume, 2.0, "f32").mu
    mode = oracle.F64X2 if precision == "f64" else oracle.F32X2
    s = oracle.OracleSession(o, volume, 2.0, mu, mode)
    s.reset(gt[0])
    poses = np.zeros((frames, 4, 4))
    tracked = np.zeros(frames, np.int8)
    t0 = time.perf_counter()
    for f in range(frames):
        s.frame(depth[f].astype(np.float32))
        poses[f] = s.pose()
        tracked[f] = s.stats()["tracked"]
        if log and f % 50 == 49:
            log("%d^3 %s frame %d: %.1f s, error vs ground truth %.4f m"
                % (volume, precision, f + 1, time.perf_counter() - t0,
                   np.linalg.norm(poses[f][:3, 3] - gt[f][:3, 3])))
    return poses, tracked


def main():
    depth, gt = synth.sequence(0, max(c[2] for c in CONFIGS.values()))
    out = {"gt": gt.astype(np.float64)}
    for name, (volume, precision, frames) in CONFIGS.items():
        poses, tracked = run(volume, precision, frames, depth, gt, log=lambda m: print(m, flush=True))
        out[name + "_poses"] = poses
        out[name + "_tracked"] = tracked
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)
    write_report(out)


def ate(P, Q):
    """Per-frame translation distance between two (F, 4, 4) trajectories."""
    return np.linalg.norm(P[:, :3, 3] - Q[:, :3, 3], axis=1)


def write_report(d, path=os.path.join(HERE, "oracle_drift.md")):
    """The oracle's drift against ground truth, and FP32 against FP64, per
    50 frames (from oracle_trajectories.npz)."""
    gt = d["gt"]
    lines = ["# The oracle's free-running frame loop against ground truth",
             "",
             "Written by tests/golden/make_trajectories.py from oracle_trajectories.npz: the CPU",
             "oracle (the reference's shaders and host code restated in C) on synthetic session 0,",
             "truncation 3 voxels, ICP 10/5/4.  Translation error in mm after frame k; `tracked`",
             "is the fraction of frames accepted so far.",
             ""]
    cols = list(CONFIGS)
    lines.append("| frame | " + " | ".join("%s err mm | tracked" % c for c in cols) + " | c2 f32 vs c3 f64 mm |")
    lines.append("|---|" + "---|---|" * len(cols) + "---|")
    last = max(CONFIGS[c][2] for c in cols)
    for k in [10, 20, 50] + list(range(100, last + 1, 100)):
        row = ["%d" % k]
        for c in cols:
            P, t = d[c + "_poses"], d[c + "_tracked"]
            if k <= len(P):
                row.append("%.1f | %.2f" % (1e3 * ate(P[k - 1:k], gt[k - 1:k])[0], t[:k].mean()))
            else:
                row.append("- | -")
        P2, P3 = d["c2_512_f32_poses"], d["c3_512_f64_poses"]
        row.append("%.2f" % (1e3 * ate(P2[k - 1:k], P3[k - 1:k])[0]) if k <= len(P2) else "-")
        lines.append("| " + " | ".join(row) + " |")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
