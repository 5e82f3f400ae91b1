"""CPU baseline with the reference's own compute shaders: gFusion's frame loop
(oracle.reference.RefSession) on Mesa's llvmpipe over the seeded synthetic
sequence of session 0, N^3 RG16I TSDF (the reference's only volume format),
2 m cube, truncation 3 voxels, ICP 10/5/4.  Prints one JSON line: frames/s
over the timed frames, milliseconds per frame by phase (each phase closed by
glFinish), the llvmpipe thread count and the host's core count.

Test infrastructure: run as its own process (bench.py starts it that way), so
Mesa's LLVM never shares a process with the GPU runtime.

    python -m oracle.ref_baseline --volume 512 --frames 20 --warm 2 --threads 16
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--frames", type=int, default=20)
    ap.add_argument("--warm", type=int, default=2)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args(argv)
    # llvmpipe sizes its thread pool when the screen is created
    os.environ["LP_NUM_THREADS"] = str(args.threads)

    import numpy as np

    from oglfusionkinectv2_amd import kinect_config, synth
    from oracle import reference as R

    total = args.warm + args.frames
    depth, gt = synth.sequence(0, total)
    mu = kinect_config(args.volume, 2.0, "f32").mu
    t_setup = time.perf_counter()
    s = R.RefSession(args.volume, 2.0, mu)
    s.reset(gt[0])
    for f in range(args.warm):
        s.frame(depth[f].astype(np.float32))
    setup_s = time.perf_counter() - t_setup
    s.times = dict.fromkeys(s.PHASES, 0.0)
    tracked = 0
    t0 = time.perf_counter()
    for f in range(args.warm, total):
        tracked += s.frame(depth[f].astype(np.float32))
    dt = time.perf_counter() - t0
    err = float(np.linalg.norm(s.pose()[:3, 3] - gt[total - 1][:3, 3]))
    print(json.dumps({
        "frames_per_s": args.frames / dt,
        "seconds": dt,
        "frames": args.frames,
        "warm": args.warm,
        "volume": args.volume,
        "ms_per_frame_by_phase": {k: 1e3 * v / args.frames for k, v in s.times.items()},
        "tracked": tracked,
        "final_translation_error_vs_gt_m": err,
        "lp_num_threads": args.threads,
        "host_cores": os.cpu_count(),
        "renderer": R.RefShaders.info(),
        "setup_and_warm_s": setup_s,
    }), flush=True)


if __name__ == "__main__":
    main()
