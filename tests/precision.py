"""Precision of the GPU path against the FP64 oracle, or against the FP32
oracle with the GPU's own storage type (the reference semantics; `against`),
as test infrastructure.

Two measurements on the same seeded sequence:

* free-running: the GPU session (FP16 or FP32 TSDF) and the FP64 oracle each
  track and fuse the frames on their own; reported are the per-frame
  translation difference (ATE against the FP64 trajectory), and at the end
  the TSDF difference over voxels both have observed and the raycast vertex
  difference over pixels both hit.  This includes the growth of differences
  through the feedback loop (track -> integrate -> raycast -> track), which
  the reference's own FP32 arithmetic shows as well (reported beside it: the
  FP32 oracle against the FP64 oracle).
* one frame from pinned state: before frame k the GPU session is loaded with
  the FP64 oracle's state (TSDF rounded to the storage type, reference maps,
  pose) and runs that one frame; reported are the pose, TSDF and raycast
  differences after it.  This is the error the precision choice adds per
  frame.

    python tests/precision.py --volume 512 --frames 30 --storage f16
"""
import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def _pose_err(P, Q):
    dt = float(np.linalg.norm(P[:3, 3].astype(np.float64) - Q[:3, 3].astype(np.float64)))
    R = P[:3, :3].astype(np.float64).T @ Q[:3, :3].astype(np.float64)
    # angle from sin (skew part) and cos (trace): arccos alone loses ~0.3 mrad
    # to float32 rounding near the identity
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    ang = float(np.arctan2(s, (np.trace(R) - 1.0) / 2.0))
    return dt, ang


def _tsdf_err(v, w, V64):
    v = v.astype(np.float64)
    w = w.astype(np.float64)
    m = (w > 0) & (V64[..., 1] > 0)
    d = np.abs(v - V64[..., 0])[m]
    if d.size == 0:
        return dict(voxels=0)
    return dict(voxels=int(d.size), mean=float(d.mean()), p99=float(np.percentile(d, 99)),
                max=float(d.max()), frac_gt_1e3=float((d > 1e-3).mean()), frac_exact=float((d == 0).mean()),
                weight_differs=float(((w > 0) != (V64[..., 1] > 0)).mean()))


def _ray_err(R, R64, voxel):
    a = np.abs(R[..., :3]).sum(-1) > 0
    b = np.abs(R64[..., :3]).sum(-1) > 0
    m = a & b
    d = np.linalg.norm(R[..., :3].astype(np.float64) - R64[..., :3].astype(np.float64), axis=-1)[m]
    if d.size == 0:
        return dict(pixels=0)
    return dict(pixels=int(d.size), median_mm=float(np.median(d) * 1e3), p99_mm=float(np.percentile(d, 99) * 1e3),
                max_mm=float(d.max() * 1e3), p99_voxels=float(np.percentile(d, 99) / voxel),
                hit_flips=float((a != b).mean()))


def evaluate(volume=512, frames=20, storage="f16", pinned=(), session=0, with_f32_oracle=True, against="f64",
             free_running=True, options=None):
    """Runs the comparison; returns a JSON-able dict.  `pinned`: frame
    indices at which the one-frame pinned comparison is made.  against:
    "f64" (the FP64 oracle, FP64 TSDF) or "f32" (the FP32 oracle with the
    GPU's own storage type, i.e. the reference semantics).  free_running:
    also run a second GPU session on its own and compare it at every frame.
    options: {kf option: value} set on the GPU sessions."""
    import torch  # noqa: F401  (initialises the device for the library)

    from oglfusionkinectv2_amd import GFusion, kinect_config, synth

    depth, _gt = synth.sequence(session, frames)
    cfg = kinect_config(volume, 2.0, storage)
    vox = float(np.float32(2.0) / np.float32(volume))
    mu = cfg.mu

    def gpu():
        g = GFusion()
        g.setConfig(kinect_config(volume, 2.0, storage))
        g.allocateBuffers()
        for k, v in (options or {}).items():
            g.setOption(k, v)
        g.Reset(_gt[0])
        return g

    g, gp = (gpu() if free_running else None), (gpu() if pinned else None)
    if against == "f64":
        s64 = oracle.OracleSession(oracle.Oracle("f64"), volume, 2.0, mu, oracle.F64X2)
    else:
        mode = {"f16": oracle.F16X2, "f32": oracle.F32X2}[storage]
        s64 = oracle.OracleSession(oracle.Oracle("f32"), volume, 2.0, mu, mode)
        with_f32_oracle = False
    s64.reset(_gt[0])
    s32 = None
    if with_f32_oracle:
        s32 = oracle.OracleSession(oracle.Oracle("f32"), volume, 2.0, mu, oracle.F32X2)
        s32.reset(_gt[0])
    plane = {"f16": np.float16, "f32": np.float32}[storage]
    ate, ate32, pin = [], [], []
    for f in range(frames):
        d = depth[f].astype(np.float32)
        if f in pinned:
            V0 = s64.volume()
            if against == "f64":
                gp.setVolume(V0[..., 0].astype(plane), V0[..., 1].astype(plane))
            else:  # the oracle's own planes in the GPU's storage type
                gp.setVolume(V0[..., 0].view(plane), V0[..., 1].view(plane))
            del V0
            gp.setReferenceMaps(s64.map(3).astype(np.float32), s64.map(4).astype(np.float32))
            gp.setPose(s64.pose().astype(np.float32))
        s64.frame(d)
        Q = s64.pose()
        if g is not None:
            g.frame(d)
            ate.append(_pose_err(g.getPose(), Q))
        if s32 is not None:
            s32.frame(d)
            ate32.append(_pose_err(s32.pose(), Q))
        if gp is not None:
            gp.frame(d)  # every frame, so its reset window and counters match the oracle's
        if f in pinned:
            dt, ang = _pose_err(gp.getPose(), Q)
            V64 = _as_values(s64.volume(), against, storage)
            v, w = gp.getVolume()
            pin.append(dict(frame=f, pose_mm=dt * 1e3, pose_mrad=ang * 1e3, tsdf=_tsdf_err(v, w, V64),
                            raycast=_ray_err(gp.getVerts(), s64.map(3), vox),
                            tracked=gp.status()["tracked"], oracle_tracked=s64.stats()["tracked"]))
            del V64
    out = dict(volume=volume, frames=frames, storage=storage, voxel_mm=vox * 1e3, against=against, pinned=pin)
    if g is not None:
        V64 = _as_values(s64.volume(), against, storage)
        v, w = g.getVolume()
        out["free_running"] = dict(ate_max_mm=max(a for a, _ in ate) * 1e3, ate_last_mm=ate[-1][0] * 1e3,
                                   rot_max_mrad=max(r for _, r in ate) * 1e3,
                                   tsdf=_tsdf_err(v, w, V64), raycast=_ray_err(g.getVerts(), s64.map(3), vox))
        del V64
    if s32 is not None:
        out["f32_oracle_vs_f64_oracle"] = dict(ate_max_mm=max(a for a, _ in ate32) * 1e3,
                                               ate_last_mm=ate32[-1][0] * 1e3)
    if g is not None:
        g.close()
    if gp is not None:
        gp.close()
    return out


def _as_values(vol, against, storage):
    """The oracle volume as float (value, weight): FP64 planes as they are,
    the FP32 oracle's FP16 planes decoded."""
    if against == "f32" and storage == "f16":
        return vol.view(np.float16).astype(np.float32)
    return vol


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--frames", type=int, default=30)
    ap.add_argument("--storage", default="f16")
    ap.add_argument("--pinned", default="5,15,25")
    ap.add_argument("--against", default="f64", choices=("f64", "f32"))
    args = ap.parse_args()
    pinned = tuple(int(x) for x in args.pinned.split(",") if x)
    print(json.dumps(evaluate(args.volume, args.frames, args.storage, pinned, against=args.against)))


if __name__ == "__main__":
    main()
