"""Profiling driver: runs the frame loop on resident synthetic depth for a
fixed number of frames (for rocprofv3 kernel traces and counters), followed by
a plain device copy of known size that calibrates the HBM byte counters.

    python -m oglfusionkinectv2_amd.profile_run --volume 512 --storage f32 --frames 40
"""
import argparse
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--frames", type=int, default=40)
    ap.add_argument("--copy-mib", type=int, default=1024)
    args = ap.parse_args()

    import torch
    from oglfusionkinectv2_amd import GFusion, kinect_config, synth
    from oglfusionkinectv2_amd import _lib as K

    L = K.lib()
    depth, poses = synth.sequence(0, args.frames)
    dev = torch.from_numpy(depth.view(np.int16)).to("cuda")
    g = GFusion()
    g.setConfig(kinect_config(args.volume, 2.0, args.storage))
    g.allocateBuffers()
    g.Reset(poses[0])
    fb = depth[0].nbytes
    for i in range(args.frames):
        K.check(L.kf_frame_u16_dev(g.session, C.c_void_p(dev.data_ptr() + i * fb)))
    g.synchronize()
    print("frames done:", args.frames, "status", g.status(), flush=True)
    n = args.copy_mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copy done: %d MiB x 5" % args.copy_mib, flush=True)


if __name__ == "__main__":
    main()
