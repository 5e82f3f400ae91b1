"""Benchmark: KinectFusion frames/s on one MI355X at a 512^3 TSDF.

Workload (the headline configuration): 512^3 FP32 TSDF + FP32 weight over a
2 m cube, truncation 3 voxels, 3-level ICP with 10/5/4 iterations (coarse to
fine), 512x424 uint16 depth from the seeded synthetic generator (session 0),
single session.  A "step" is one frame of the kinect.cpp:654-695 loop:
pyramid, Track (device-side ICP), gated integrate, raycast.  All depth frames
are resident in HBM before the timed region.

Also reported: the same loop with the FP16 TSDF (4 B/voxel), per-frame
p50/p99 latency, the dominant kernel's roofline figures from HIP events
recorded on the session stream around it inside the timed region, and the CPU
oracle timed on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=2)
    return ap.parse_args()


def make_session(volume, storage, stream):
    from oglfusionkinectv2_amd import GFusion, kinect_config
    g = GFusion(stream=stream)
    g.setConfig(kinect_config(volume, 2.0, storage))
    g.allocateBuffers()
    return g


def run_loop(torch, L, g, frames_dev, poses, warmup, steps, stream, time_kernels):
    """Runs warmup + steps frames; returns (elapsed_s, per-frame ms list,
    per-kernel mean ms dict)."""
    import ctypes as C
    from oglfusionkinectv2_amd import _lib as K
    s = g.session
    fb = frames_dev[0].numel() * 2
    base = frames_dev.data_ptr()
    g.Reset(poses[0])
    g.synchronize()
    ev = torch.cuda.Event

    def frame(i, marks):
        ptr = C.c_void_p(base + i * fb)
        if marks is not None:
            marks[0].record(stream)
        K.check(L.kf_depth_to_vertex_u16_dev(s, ptr))
        if marks is not None:
            marks[1].record(stream)
        K.check(L.kf_track_frame(s))
        if marks is not None:
            marks[2].record(stream)
        K.check(L.kf_integrate(s, K.INTEGRATE_GATED))
        if marks is not None:
            marks[3].record(stream)
        K.check(L.kf_raycast(s))
        if marks is not None:
            marks[4].record(stream)

    for i in range(warmup):
        frame(i, None)
    torch.cuda.synchronize()
    marks = [[ev(enable_timing=True) for _ in range(5)] for _ in range(steps)] if time_kernels else None
    t0 = time.perf_counter()
    for i in range(steps):
        frame(warmup + i, marks[i] if marks else None)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per_frame, kern = [], {}
    if marks:
        names = ["pyramid", "icp", "integrate", "raycast"]
        acc = {n: [] for n in names}
        for m in marks:
            per_frame.append(m[0].elapsed_time(m[4]))
            for k, n in enumerate(names):
                acc[n].append(m[k].elapsed_time(m[k + 1]))
        kern = {n: float(np.mean(v)) for n, v in acc.items()}
    return elapsed, per_frame, kern


def cpu_baseline(depth, poses, volume, frames):
    """The FP32 C oracle on the same workload, `frames` frames after reset."""
    import oracle
    o = oracle.Oracle("f32")
    vox = np.float32(2.0) / np.float32(volume)
    mu = float(vox * np.float32(3.0))
    s = oracle.OracleSession(o, volume, 2.0, mu, oracle.F32X2)
    s.reset(poses[0])
    t0 = time.perf_counter()
    for f in range(frames):
        s.frame(depth[f].astype(np.float32))
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return frames / dt, threads, dt


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (sessions shard "
              "across GPUs with no communication, see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from oglfusionkinectv2_amd import _lib as K
    from oglfusionkinectv2_amd import synth

    L = K.lib()
    nframes = args.warmup + args.steps
    t_gen = time.perf_counter()
    depth, poses = synth.sequence(0, nframes)
    t_gen = time.perf_counter() - t_gen
    frames_dev = torch.from_numpy(depth.view(np.int16)).to("cuda")
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    # --- FP32 (headline) ---------------------------------------------------
    g = make_session(args.volume, "f32", stream.cuda_stream)
    elapsed, per_frame, kern = run_loop(torch, L, g, frames_dev, poses, args.warmup, args.steps, stream, True)
    fps = args.steps / elapsed
    st = g.status()
    est = g.getPose()
    ate_gt = float(np.linalg.norm(est[:3, 3] - poses[nframes - 1][:3, 3]))
    # integrate traffic under the reference semantics, counted on the last frame
    vr, vw = g.countIntegrateTraffic()
    bv = 8
    int_bytes = (vr + vw) * bv + 424 * 512 * 4
    dom = max(kern, key=kern.get)
    g.close()

    # --- FP16 --------------------------------------------------------------
    fps16, kern16 = None, {}
    if not args.no_fp16:
        g16 = make_session(args.volume, "f16", stream.cuda_stream)
        e16, pf16, kern16 = run_loop(torch, L, g16, frames_dev, poses, args.warmup, args.steps, stream, True)
        fps16 = args.steps / e16
        g16.close()

    # --- CPU baseline --------------------------------------------------------
    cpu = None
    if not args.no_cpu:
        cfps, threads, cdt = cpu_baseline(depth, poses, args.volume, args.cpu_frames)
        cpu = {"value": cfps, "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": "frames 0-%d of the same %d^3 FP32 sequence (after reset), C oracle, "
                         "%d OpenMP threads, %.1f s" % (args.cpu_frames - 1, args.volume, threads, cdt)}

    int_ms = kern["integrate"]
    roof = {
        "kernel": "k_integrate<StoreF32>",
        "bound": "hbm",
        "achieved": int_bytes / (int_ms * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": int_bytes / (int_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": None,
        "algorithmic_bytes": int(int_bytes),
        "avg_ms": int_ms,
    }
    pf = np.array(per_frame)
    out = {
        "metric": "frames/s per GPU at 512^3 TSDF (FP32 TSDF+FP32 weight, single session)",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 / fps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic room + spheres + boxes, Kinect v2 noise model)",
        "config": {"workload": "plain KinectFusion loop, %d^3 FP32 TSDF over 2 m, 512x424 depth, ICP 10/5/4, "
                               "single session" % args.volume, "volume": args.volume, "frames": nframes},
        "fp16_frames_per_s": fps16,
        "fp16_vs_fp32": (fps16 / fps) if fps16 else None,
        "voxel_updates_per_s": fps * args.volume ** 3,
        "latency_ms": {"p50": float(np.percentile(pf, 50)), "p99": float(np.percentile(pf, 99))},
        "kernel_ms": kern,
        "kernel_ms_fp16": kern16,
        "dominant_kernel": dom,
        "roofline": roof,
        "cpu_baseline": cpu,
        "tracked_last": st["tracked"],
        "final_translation_error_vs_gt_m": ate_gt,
        "generator_s": t_gen,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
