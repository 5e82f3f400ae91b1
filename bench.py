"""Benchmark: KinectFusion frames/s on one MI355X at a 512^3 TSDF.

Workload (the headline configuration): 512^3 FP32 TSDF + FP32 weight over a
2 m cube, truncation 3 voxels, 3-level ICP with 10/5/4 iterations (coarse to
fine), 512x424 uint16 depth from the seeded synthetic generator (session 0),
single session.  A "step" is one frame of the kinect.cpp:654-695 loop:
pyramid, Track (device-side ICP), gated integrate, raycast.  All depth frames
are resident in HBM before the timed region.

Also reported: the same loop with the FP16 TSDF (4 B/voxel), per-frame
p50/p99 latency, the dominant kernel's roofline figures from HIP events
carried by its dispatch packets on the session stream, the CPU
oracle timed on a bounded sample of the same workload, and the multi-session
FP16 legs (saturated 16 x 512^3 batch, the uneven 16-session mix, 16 x 512^3
served at 30 Hz, Poisson arrivals under a 200 GB budget) under
"multi_session".
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
    ap.add_argument("--cpu-frames", type=int, default=60)
    ap.add_argument("--fast-math", type=int, default=1)
    ap.add_argument("--no-serve", action="store_true", help="skip the batched and served legs")
    ap.add_argument("--batch-ticks", type=int, default=100)
    ap.add_argument("--serve-seconds", type=float, default=8.0,
                    help="length of the fixed 16 x 512^3 at 30 Hz served run")
    ap.add_argument("--poisson-seconds", type=float, default=40.0,
                    help="how much of the Poisson arrival process (64 sessions at 1/s) to serve")
    return ap.parse_args()


def make_session(volume, storage, stream, fast_math):
    from oglfusionkinectv2_amd import GFusion, kinect_config
    from oglfusionkinectv2_amd import _lib as K
    g = GFusion(stream=stream)
    g.setConfig(kinect_config(volume, 2.0, storage))
    g.allocateBuffers()
    g.setOption(K.OPT_FAST_MATH, 1 if fast_math else 0)
    return g


def run_frames(torch, L, g, frames_dev, poses, warmup, steps, stream):
    """Headline pass: the product path (kf_frame_u16_dev, one pyramid launch +
    one graph replay per frame).  Returns elapsed seconds and per-frame
    latencies from events recorded between frames on the session stream."""
    import ctypes as C
    from oglfusionkinectv2_amd import _lib as K
    s = g.session
    fb = frames_dev[0].numel() * 2
    base = frames_dev.data_ptr()
    g.Reset(poses[0])
    for i in range(warmup):
        K.check(L.kf_frame_u16_dev(s, C.c_void_p(base + i * fb)))
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(steps):
        K.check(L.kf_frame_u16_dev(s, C.c_void_p(base + (warmup + i) * fb)))
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    lat = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
    return elapsed, lat


def run_kernels(torch, L, g, frames_dev, poses, warmup, steps, count_every=25):
    """Per-kernel pass: the same frames with stage timing on (the pyramid,
    integrate and raycast launches carry start / stop events in their own
    dispatch packets, so each figure is that kernel's duration; ICP is its
    first iteration's start to the finish kernel's end).  Every `count_every`
    frames the integrate and raycast traffic of that frame is counted
    (separate passes after the frame): integrate under the reference semantics
    and as the default kernel moves it, the raycast's distinct bytes."""
    import ctypes as C
    from oglfusionkinectv2_amd import _lib as K
    s = g.session
    fb = frames_dev[0].numel() * 2
    base = frames_dev.data_ptr()
    g.setOption(K.OPT_STAGE_TIMING, 1)
    g.Reset(poses[0])
    for i in range(warmup):
        K.check(L.kf_frame_u16_dev(s, C.c_void_p(base + i * fb)))
    torch.cuda.synchronize()
    acc = {n: [] for n in ("pyramid", "icp", "integrate", "raycast")}
    counts, ray_counts = [], []
    for i in range(steps):
        K.check(L.kf_frame_u16_dev(s, C.c_void_p(base + (warmup + i) * fb)))
        for n, v in g.stageTimes().items():
            acc[n].append(v)
        if i % count_every == count_every - 1:
            counts.append(g.countIntegrateTraffic() + g.countIntegrateFastTraffic())
            ray_counts.append(g.countRaycastTraffic())
    g.setOption(K.OPT_STAGE_TIMING, 0)
    return {n: float(np.mean(v)) for n, v in acc.items()}, counts, ray_counts


def row_timings(torch, L, g, frames_dev, nframes, reps=3):
    """The paths beside the frame loop on the same 512^3 session after the
    timed frames: TrackSDF (one persistent launch, from the session's
    m_pose_eig), recoverPose's no-op launch when tracking succeeded (the
    frame graph carries it), the surface export (marching cubes over the
    whole TSDF) and intensityProjection.  Wall-clock over synchronised
    calls, best of `reps`."""
    import ctypes as C
    from oglfusionkinectv2_amd import _lib as K
    from oglfusionkinectv2_amd import MCubes
    s = g.session
    out = {}

    def best(fn):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = fn()
            K.check(L.kf_synchronize(s))
            ts.append(time.perf_counter() - t0)
        return 1e3 * min(ts), r

    g.setPoseEig(g.getPose())
    out["track_sdf_ms"], tracked = best(g.TrackSDF)
    out["track_sdf_tracked"] = bool(tracked)
    out["track_sdf_iterations"] = g.sdfState()["iterations"]
    mc = MCubes()
    mc.setVolumeTexture(g)
    mc.setIsolevel(0.0)
    out["mesh_export_ms"], ntri = best(mc.generateMarchingCubes)
    out["mesh_triangles"] = int(ntri)
    out["intensity_projection_ms"], _ = best(lambda: K.check(L.kf_intensity_projection(s)))
    return out


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "frame_loop_512_f32_hbm_traffic.md")


def measured_traffic():
    """HBM bytes per launch of each kernel at 512^3 FP32 from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (lines `traffic_bytes_per_launch
    <kernel> <bytes>`, already scaled by the copy calibration)."""
    out = {}
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                p = line.split()
                if len(p) == 3 and p[0] == "traffic_bytes_per_launch":
                    out[p[1]] = int(p[2])
    except OSError:
        pass
    return out


def roofline(kernel, nbytes, ms, traffic):
    """HBM roofline of one kernel: algorithmic bytes per launch over its mean
    launch duration (HIP events in the dispatch packets on the session stream)."""
    ach = nbytes / (ms * 1e-3) / 1e9
    key = kernel.replace(" ", "")
    return {"kernel": kernel, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": traffic.get(key), "algorithmic_bytes": int(nbytes),
            "avg_ms": ms, "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if key in traffic else None}


def cpu_baseline(depth, poses, volume, frames, warm=2):
    """The FP32 C oracle on the same workload: `warm` untimed frames after
    reset, then `frames` timed frames."""
    import oracle
    o = oracle.Oracle("f32")
    vox = np.float32(2.0) / np.float32(volume)
    mu = float(vox * np.float32(3.0))
    s = oracle.OracleSession(o, volume, 2.0, mu, oracle.F32X2)
    s.reset(poses[0])
    for f in range(warm):
        s.frame(depth[f].astype(np.float32))
    t0 = time.perf_counter()
    for f in range(warm, warm + frames):
        s.frame(depth[f].astype(np.float32))
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return frames / dt, threads, dt


def serving_legs(torch, args, fps32):
    """Multi-session legs, all FP16 TSDF:
    - saturated: 16 x 512^3 sessions with a frame due every tick (offline
      throughput of the batched path), compared with the FP32 single session;
    - batched: the uneven mix 4 x 128^3, 6 x 256^3, 4 x 384^3, 2 x 512^3, each
      session its own seeded scene, every session due every tick;
    - served: 16 x 512^3 sessions at 30 Hz, first frames spread over one
      period, real-time request loop with continuous batching;
    - poisson: sessions arriving at 1/s (64 in all), sizes {256..768}^3, 15 or
      30 Hz, 200-900 frames each, 200 GB HBM budget; the first
      `poisson_seconds` of the process are served."""
    from oglfusionkinectv2_amd import serve
    out = {}
    sat = serve.saturate(torch, [512] * 16, args.batch_ticks, sequences=16,
                         pool_frames=args.batch_ticks + 10)
    sat["vs_fp32_single_session"] = sat["frames_per_s"] / fps32
    out["saturated_16x512_f16"] = sat
    mix = [128] * 4 + [256] * 6 + [384] * 4 + [512] * 2
    out["batched_mix_f16"] = serve.saturate(torch, mix, args.batch_ticks, sequences=len(mix),
                                            pool_frames=args.batch_ticks + 10)
    out["served_16x512_30hz_f16"] = serve.serve(torch, serve.fixed_sessions(16, 512, 30.0, args.serve_seconds),
                                                args.serve_seconds, sequences=16)
    if args.poisson_seconds > 0:
        out["served_poisson_f16"] = serve.serve(torch, serve.poisson_sessions(), args.poisson_seconds,
                                                sequences=16)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this benchmark runs on one GPU; --gpus %d is not supported (sessions are "
              "independent and shard across GPUs with no communication)" % args.gpus, file=sys.stderr)
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
    g = make_session(args.volume, "f32", stream.cuda_stream, args.fast_math)
    elapsed, lat = run_frames(torch, L, g, frames_dev, poses, args.warmup, args.steps, stream)
    fps = args.steps / elapsed
    st = g.status()
    est = g.getPose()
    ate_gt = float(np.linalg.norm(est[:3, 3] - poses[nframes - 1][:3, 3]))
    kern, counts, ray_counts = run_kernels(torch, L, g, frames_dev, poses, args.warmup, args.steps)
    # integrate traffic (sampled frames): under the reference semantics each
    # read voxel moves value + weight in and out, each written-only voxel out;
    # the default kernel skips chunks that stay (0, 0) and stores of (0, 0)
    # over (0, 0), so its algorithmic bytes are the voxels it reads plus the
    # voxels it writes (value + weight each); both plus the depth image
    bv = 8
    vr = float(np.mean([c[0] for c in counts])) if counts else 0.0
    vw = float(np.mean([c[1] for c in counts])) if counts else 0.0
    fr = float(np.mean([c[2] for c in counts])) if counts else 0.0
    fw = float(np.mean([c[3] for c in counts])) if counts else 0.0
    int_ref_bytes = (vr + vw) * bv + 424 * 512 * 4
    int_bytes = (fr + fw) * bv + 424 * 512 * 4
    # raycast: distinct voxel values (4 B) and occupancy-flag bytes read, plus
    # the vertex and normal maps written (2 x float4 per pixel)
    rv = float(np.mean([c[0] for c in ray_counts])) if ray_counts else 0.0
    rf = float(np.mean([c[1] for c in ray_counts])) if ray_counts else 0.0
    ray_bytes = rv * 4 + rf + 424 * 512 * 32
    dom = max(kern, key=kern.get)
    extras = row_timings(torch, L, g, frames_dev, args.warmup + args.steps)
    g.close()

    # --- FP16 --------------------------------------------------------------
    fps16, kern16, lat16 = None, {}, None
    if not args.no_fp16:
        g16 = make_session(args.volume, "f16", stream.cuda_stream, args.fast_math)
        e16, lat16 = run_frames(torch, L, g16, frames_dev, poses, args.warmup, args.steps, stream)
        fps16 = args.steps / e16
        kern16, _, _ = run_kernels(torch, L, g16, frames_dev, poses, args.warmup, args.steps)
        g16.close()

    # --- multi-session -----------------------------------------------------
    multi = None
    if not args.no_serve:
        multi = serving_legs(torch, args, fps)

    # --- CPU baseline --------------------------------------------------------
    cpu = None
    if not args.no_cpu:
        cfps, threads, cdt = cpu_baseline(depth, poses, args.volume, args.cpu_frames)
        cpu = {"value": cfps, "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": "frames 2-%d of the same %d^3 FP32 sequence (frames 0-1 untimed), C oracle, "
                         "%d OpenMP threads, %.1f s" % (args.cpu_frames + 1, args.volume, threads, cdt)}

    traffic = measured_traffic() if args.volume == 512 else {}
    rooflines = {
        "raycast": roofline("k_raycast<StoreF32>", ray_bytes, kern["raycast"], traffic),
        "integrate": roofline("k_integrate_fast<StoreF32, true>", int_bytes, kern["integrate"], traffic),
    }
    rooflines["integrate"]["reference_semantics_bytes"] = int(int_ref_bytes)
    rooflines["integrate"]["reference_semantics_GBps"] = int_ref_bytes / (kern["integrate"] * 1e-3) / 1e9
    roof = rooflines.get(dom) or rooflines["raycast"]
    pf = np.array(lat)
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
        "latency_ms_fp16": ({"p50": float(np.percentile(lat16, 50)), "p99": float(np.percentile(lat16, 99))}
                            if lat16 else None),
        "fast_math": args.fast_math,
        "dominant_kernel": dom,
        "roofline": roof,
        "roofline_by_kernel": rooflines,
        "cpu_baseline": cpu,
        "beside_the_loop": extras,
        "multi_session": multi,
        "tracked_last": st["tracked"],
        "final_translation_error_vs_gt_m": ate_gt,
        "generator_s": t_gen,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
