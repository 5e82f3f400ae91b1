"""Benchmark: KinectFusion frames/s on one MI355X at a 512^3 TSDF.

Workload (the headline configuration): 512^3 FP32 TSDF + FP32 weight over a
2 m cube, truncation 3 voxels, 3-level ICP with 10/5/4 iterations (coarse to
fine), 512x424 uint16 depth from the seeded synthetic generator (session 0),
single session.  A "step" is one frame of the kinect.cpp:654-695 loop:
pyramid, Track (device-side ICP), gated integrate, raycast.  All depth frames
are resident in HBM before the timed region.

Also reported: the same loop with the FP16 TSDF (4 B/voxel), per-frame
p50/p99 latency, the dominant kernel's roofline figures from HIP events
carried by its dispatch packets on the session stream, the tracked fraction
and translation error against the ground-truth trajectory, the frame rate
with the per-frame host upload (kf_frame_f32), the CPU baseline on bounded
samples of the same workload (the reference's own compute shaders on Mesa
llvmpipe, in a child process, and the C oracle), and the multi-session legs (saturated
16 x 512^3 batch in FP16 and FP32, the uneven 16-session mix, 16 and 64 x
512^3 served at 30 Hz, Poisson arrivals under a 200 GB budget) under
"multi_session".  Every leg after the headline records an "error" field
instead of stopping the run, so the JSON line is always printed.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=60)
    ap.add_argument("--cpu-warm", type=int, default=2, help="untimed CPU oracle frames after its reset")
    ap.add_argument("--ref-frames", type=int, default=20,
                    help="timed frames of the reference's shaders on llvmpipe (0: the C oracle alone)")
    ap.add_argument("--fast-math", type=int, default=1)
    ap.add_argument("--no-serve", action="store_true", help="skip the batched and served legs")
    ap.add_argument("--batch-ticks", type=int, default=100)
    ap.add_argument("--serve-seconds", type=float, default=8.0,
                    help="length of each fixed-membership served run (16 and 64 x 512^3 at 30 Hz)")
    ap.add_argument("--poisson-seconds", type=float, default=40.0,
                    help="how much of the Poisson arrival process (64 sessions at 1/s) to serve")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the 600-frame accuracy leg")
    ap.add_argument("--plan", action="store_true",
                    help="print the frame plan for these arguments and exit (no GPU)")
    return ap.parse_args(argv)


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


def run_kernels(torch, L, g, frames_dev, poses, warmup, steps, count_every=25, count=True):
    """Per-kernel pass: the same frames with stage timing on (the pyramid,
    integrate and raycast launches carry start / stop events in their own
    dispatch packets, so each figure is that kernel's duration; ICP is its
    first iteration's start to the finish kernel's end).  After the frames of
    count_schedule() the integrate and raycast traffic of that frame is
    counted (separate passes after the frame): integrate under the reference
    semantics and as the default kernel moves it, the raycast's distinct
    bytes.  Every frame's tracking decision and pose are read back, giving
    the tracked fraction and the translation error against the ground-truth
    trajectory (the frames are the headline pass's, from the same reset, so
    the poses are the same)."""
    import ctypes as C
    from oglfusionkinectv2_amd import _lib as K
    s = g.session
    fb = frames_dev[0].numel() * 2
    base = frames_dev.data_ptr()
    g.setOption(K.OPT_STAGE_TIMING, 1)
    g.Reset(poses[0])
    tracked, err = [], []

    def record(f):
        tracked.append(g.status()["tracked"])
        err.append(float(np.linalg.norm(g.getPose()[:3, 3] - poses[f][:3, 3])))

    for i in range(warmup):
        K.check(L.kf_frame_u16_dev(s, C.c_void_p(base + i * fb)))
        record(i)
    torch.cuda.synchronize()
    acc = {n: [] for n in ("pyramid", "icp", "integrate", "raycast")}
    counts, ray_counts = [], []
    sched = set(count_schedule(steps, count_every)) if count else set()
    for i in range(steps):
        K.check(L.kf_frame_u16_dev(s, C.c_void_p(base + (warmup + i) * fb)))
        for n, v in g.stageTimes().items():
            acc[n].append(v)
        record(warmup + i)
        if i in sched:
            counts.append(g.countIntegrateTraffic() + g.countIntegrateFastTraffic())
            ray_counts.append(g.countRaycastTraffic())
    g.setOption(K.OPT_STAGE_TIMING, 0)
    e = np.array(err)
    t = np.array(tracked)
    timed = slice(warmup, warmup + steps)
    accuracy = {
        "frames": len(e),
        "tracked_fraction_timed": float(t[timed].mean()) if steps else None,
        "tracked_fraction_all": float(t.mean()),
        "translation_error_vs_gt_m": {"timed_mean": float(e[timed].mean()) if steps else None,
                                      "timed_max": float(e[timed].max()) if steps else None,
                                      "ate_rmse_all": float(np.sqrt(np.mean(e ** 2))),
                                      "final": float(e[-1])},
        "first_frame_over_1cm": int(np.argmax(e > 0.01)) if (e > 0.01).any() else None,
    }
    return {n: float(np.mean(v)) for n, v in acc.items()}, counts, ray_counts, accuracy


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
    reset, then `frames` timed frames (depth must hold warm + frames)."""
    import oracle
    if len(depth) < warm + frames:
        raise ValueError("cpu_baseline needs %d frames, got %d" % (warm + frames, len(depth)))
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
    """Multi-session legs (FP16 TSDF unless named _f32):
    - saturated: 16 x 512^3 sessions with a frame due every tick (offline
      throughput of the batched path), FP16 and FP32, compared with the FP32
      single session;
    - batched: the uneven mix 4 x 128^3, 6 x 256^3, 4 x 384^3, 2 x 512^3, each
      session its own seeded scene, every session due every tick;
    - served: 16 and 64 x 512^3 sessions at 30 Hz (480 and 1920 frames/s
      offered), first frames spread over one period, real-time request loop
      with continuous batching;
    - poisson: sessions arriving at 1/s (64 in all), sizes {256..768}^3, 15 or
      30 Hz, 200-900 frames each, 200 GB HBM budget; the first
      `poisson_seconds` of the process are served.
    Each leg that fails records its error and the others still run."""
    from oglfusionkinectv2_amd import serve
    out = {}

    def sat(volumes, storage):
        r = serve.saturate(torch, volumes, args.batch_ticks, sequences=len(volumes), storage=storage,
                           pool_frames=args.batch_ticks + 10)
        r["vs_fp32_single_session"] = r["frames_per_s"] / fps32
        return r

    run_leg(out, "saturated_16x512_f16", lambda: sat([512] * 16, "f16"))
    run_leg(out, "saturated_16x512_f32", lambda: sat([512] * 16, "f32"))
    mix = [128] * 4 + [256] * 6 + [384] * 4 + [512] * 2
    run_leg(out, "batched_mix_f16", lambda: sat(mix, "f16"))
    run_leg(out, "served_16x512_30hz_f16",
            lambda: serve.serve(torch, serve.fixed_sessions(16, 512, 30.0, args.serve_seconds),
                                args.serve_seconds, sequences=16))
    run_leg(out, "served_64x512_30hz_f16",
            lambda: serve.serve(torch, serve.fixed_sessions(64, 512, 30.0, args.serve_seconds),
                                args.serve_seconds, sequences=16))
    if args.poisson_seconds > 0:
        run_leg(out, "served_poisson_f16",
                lambda: serve.serve(torch, serve.poisson_sessions(), args.poisson_seconds, sequences=16))
    return out


ORACLE_TRAJ = os.path.join(ROOT, "tests", "golden", "oracle_trajectories.npz")


def free_run_poses(torch, L, volume, storage, frames_dev, poses, n, options=None):
    """(n, 4, 4) poses and tracked flags of a free-running session over the
    first n resident frames (read back after every frame)."""
    import ctypes as C
    from oglfusionkinectv2_amd import _lib as K
    g = make_session(volume, storage, None, 1)
    for k, v in (options or {}).items():
        g.setOption(k, v)
    g.Reset(poses[0])
    fb = frames_dev[0].numel() * 2
    out, tracked = np.zeros((n, 4, 4)), np.zeros(n, np.int8)
    for f in range(n):
        K.check(L.kf_frame_u16_dev(g.session, C.c_void_p(frames_dev.data_ptr() + f * fb)))
        out[f] = g.getPose()
        tracked[f] = g.status()["tracked"]
    g.close()
    return out, tracked


def accuracy_leg(torch, L, volume):
    """The 600-frame sequences of the 512^3 configurations, free-running on
    the GPU, against the oracle trajectories (tests/golden/
    oracle_trajectories.npz): ATE (RMS and largest translation difference)
    against the FP64 oracle (the precision variant's reference) and the FP32
    oracle (the reference semantics), and against ground truth."""
    if volume != 512 or not os.path.exists(ORACLE_TRAJ):
        return {"skipped": "needs 512^3 and %s" % os.path.relpath(ORACLE_TRAJ, ROOT)}
    from oglfusionkinectv2_amd import _lib as K
    from oglfusionkinectv2_amd import synth
    ref = np.load(ORACLE_TRAJ)
    n = len(ref["c3_512_f64_poses"])
    depth, gt = synth.sequence(0, n)
    dev = torch.from_numpy(depth.view(np.int16)).to("cuda")

    def ate(P, Q):
        d = np.linalg.norm(P[:, :3, 3] - Q[:, :3, 3], axis=1) * 1e3
        return {"rms_mm": float(np.sqrt(np.mean(d ** 2))), "max_mm": float(d.max()),
                "frame_20_mm": float(d[min(19, len(d) - 1)])}

    out = {"frames": n, "oracle_f32_vs_f64": ate(ref["c2_512_f32_poses"], ref["c3_512_f64_poses"]),
           "oracle_f64_vs_ground_truth": ate(ref["c3_512_f64_poses"], gt)}
    legs = {"f16_fp64_pose": ("f16", {K.OPT_POSE_FP64: 1}), "f16": ("f16", {}), "f32": ("f32", {})}
    for name, (st, opts) in legs.items():
        P, t = free_run_poses(torch, L, volume, st, dev, gt, n, opts)
        out[name] = {"vs_fp64_oracle": ate(P, ref["c3_512_f64_poses"]),
                     "vs_fp32_oracle": ate(P, ref["c2_512_f32_poses"]),
                     "vs_ground_truth": ate(P, gt), "tracked_fraction": float(t.mean())}
    return out


def frame_plan(warmup, steps, cpu_frames, cpu_warm=2, with_cpu=True):
    """Frames of the seeded sequence each leg reads: the GPU legs run frames
    0 .. warmup + steps - 1; the CPU oracle runs `cpu_warm` untimed frames
    after its reset, then `cpu_frames` timed ones.  One sequence long enough
    for both is generated."""
    n_gpu = warmup + steps
    n_cpu = (cpu_warm + cpu_frames) if with_cpu and cpu_frames > 0 else 0
    return {"gpu_frames": n_gpu, "cpu_frames": n_cpu, "generate": max(n_gpu, n_cpu)}


def count_schedule(steps, every=25):
    """Frames (0-based within the timed steps) after which the integrate and
    raycast traffic is counted: every `every`-th frame, and at least the last
    one, so a short run still has a sample."""
    idx = [i for i in range(steps) if i % every == every - 1]
    if not idx and steps > 0:
        idx = [steps - 1]
    return idx


def run_leg(out, name, fn):
    """Runs one optional leg; a failure is recorded as {"error": ...} and the
    benchmark carries on, so the JSON line is always printed."""
    t0 = time.perf_counter()
    try:
        out[name] = fn()
    except Exception as e:  # noqa: BLE001 -- reported in the JSON line
        import traceback
        traceback.print_exc(file=sys.stderr)
        out[name] = {"error": "%s: %s" % (type(e).__name__, e)}
    print("bench: %s done in %.1f s" % (name, time.perf_counter() - t0), file=sys.stderr, flush=True)
    return out[name]


def run_with_upload(torch, L, g, depth_f32, poses, warmup, steps):
    """The boundary's host-buffer path (kf_frame_f32, what gFusion::
    depthToVertex(float*) takes): each frame's float mm image is copied from
    pageable host memory before the pyramid.  Returns frames/s."""
    from oglfusionkinectv2_amd import _lib as K
    s = g.session
    g.Reset(poses[0])
    for i in range(warmup):
        K.check(L.kf_frame_f32(s, depth_f32[i].ctypes.data))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        K.check(L.kf_frame_f32(s, depth_f32[warmup + i].ctypes.data))
    torch.cuda.synchronize()
    return steps / (time.perf_counter() - t0)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this benchmark runs on one GPU; --gpus %d is not supported (sessions are "
              "independent and shard across GPUs with no communication)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    plan = frame_plan(args.warmup, args.steps, args.cpu_frames, args.cpu_warm, not args.no_cpu)
    if args.plan:
        print(json.dumps({"plan": plan, "count_frames": count_schedule(args.steps)}))
        return
    import torch
    from oglfusionkinectv2_amd import _lib as K
    from oglfusionkinectv2_amd import synth

    L = K.lib()
    nframes = plan["gpu_frames"]
    t_gen = time.perf_counter()
    depth, poses = synth.sequence(0, plan["generate"])
    t_gen = time.perf_counter() - t_gen
    frames_dev = torch.from_numpy(depth[:nframes].view(np.int16)).to("cuda")
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()

    # --- FP32 (headline) ---------------------------------------------------
    g = make_session(args.volume, "f32", stream.cuda_stream, args.fast_math)
    elapsed, lat = run_frames(torch, L, g, frames_dev, poses, args.warmup, args.steps, stream)
    fps = args.steps / elapsed
    print("bench: headline %.1f frames/s" % fps, file=sys.stderr, flush=True)
    legs = {}
    kr = run_leg(legs, "kernels", lambda: run_kernels(torch, L, g, frames_dev, poses, args.warmup, args.steps))
    kern, counts, ray_counts, acc = (kr if isinstance(kr, tuple) else ({}, [], [], {}))
    # the raycast kernel the frame loop launches: the hybrid schedule (heavy
    # tiles four lanes per ray) when on, else the one-lane march
    ray_kernel = ("k_raycast_hybrid<StoreF32, 1>" if g.getOption(K.OPT_RAYCAST_HYBRID)
                  and g.getOption(K.OPT_RAYCAST_LANES) == 1 else "k_raycast<StoreF32, 1>")
    run_leg(legs, "beside_the_loop", lambda: row_timings(torch, L, g, frames_dev, nframes))
    g.close()
    depth_f32 = [depth[i].astype(np.float32) for i in range(nframes)]
    run_leg(legs, "with_upload", lambda: with_upload_leg(torch, L, args, stream, depth_f32, poses))
    del depth_f32

    # --- FP16 --------------------------------------------------------------
    fp16 = {}
    if not args.no_fp16:
        fp16 = run_leg(legs, "fp16", lambda: fp16_leg(torch, L, args, stream, frames_dev, poses))

    # --- accuracy of the 600-frame configurations against the oracle --------
    if not args.no_accuracy:
        run_leg(legs, "accuracy_600", lambda: accuracy_leg(torch, L, args.volume))

    # --- CPU baseline (before the long serving legs) -------------------------
    cpu = None
    if not args.no_cpu:
        cpu = run_leg(legs, "cpu_baseline", lambda: cpu_leg(depth, poses, args))

    # --- multi-session -----------------------------------------------------
    multi = None
    if not args.no_serve:
        multi = serving_legs(torch, args, fps)

    traffic = measured_traffic() if args.volume == 512 else {}
    rooflines, roof, dom = {}, None, None
    if kern:
        # integrate traffic (sampled frames): under the reference semantics each
        # read voxel moves value + weight in and out, each written-only voxel
        # out; the default kernel skips chunks that stay (0, 0) and stores of
        # (0, 0) over (0, 0), so its algorithmic bytes are the voxels it reads
        # plus the voxels it writes (value + weight each); both plus the depth
        bv = 8
        vr = float(np.mean([c[0] for c in counts])) if counts else 0.0
        vw = float(np.mean([c[1] for c in counts])) if counts else 0.0
        fr = float(np.mean([c[2] for c in counts])) if counts else 0.0
        fw = float(np.mean([c[3] for c in counts])) if counts else 0.0
        int_ref_bytes = (vr + vw) * bv + 424 * 512 * 4
        int_bytes = (fr + fw) * bv + 424 * 512 * 4
        # raycast: distinct voxel values (4 B) and occupancy-flag bytes read,
        # plus the vertex and normal maps written (2 x float4 per pixel)
        rv = float(np.mean([c[0] for c in ray_counts])) if ray_counts else 0.0
        rf = float(np.mean([c[1] for c in ray_counts])) if ray_counts else 0.0
        ray_bytes = rv * 4 + rf + 424 * 512 * 32
        dom = max(kern, key=kern.get)
        rooflines = {
            "raycast": roofline(ray_kernel, ray_bytes, kern["raycast"], traffic),
            "integrate": roofline("k_integrate_fast<StoreF32, true>", int_bytes, kern["integrate"], traffic),
        }
        rooflines["integrate"]["reference_semantics_bytes"] = int(int_ref_bytes)
        rooflines["integrate"]["reference_semantics_GBps"] = int_ref_bytes / (kern["integrate"] * 1e-3) / 1e9
        rooflines["raycast"]["samples"] = len(ray_counts)
        rooflines["integrate"]["samples"] = len(counts)
        roof = rooflines.get(dom) or rooflines["raycast"]
    pf = np.array(lat)
    fps16 = fp16.get("frames_per_s") if isinstance(fp16, dict) else None
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
        "fast_math": args.fast_math,
        "dominant_kernel": dom,
        "roofline": roof,
        "roofline_by_kernel": rooflines,
        "cpu_baseline": cpu,
        "with_upload_frames_per_s": legs.get("with_upload"),
        "fp16": fp16,
        "accuracy": acc,
        "accuracy_600_frames": legs.get("accuracy_600"),
        "beside_the_loop": legs.get("beside_the_loop"),
        "multi_session": multi,
        "errors": {k: v["error"] for k, v in legs.items() if isinstance(v, dict) and "error" in v} or None,
        "generator_s": t_gen,
    }
    print(json.dumps(out), flush=True)


def with_upload_leg(torch, L, args, stream, depth_f32, poses):
    g = make_session(args.volume, "f32", stream.cuda_stream, args.fast_math)
    try:
        return run_with_upload(torch, L, g, depth_f32, poses, args.warmup, args.steps)
    finally:
        g.close()


def fp16_leg(torch, L, args, stream, frames_dev, poses):
    g = make_session(args.volume, "f16", stream.cuda_stream, args.fast_math)
    try:
        e16, lat16 = run_frames(torch, L, g, frames_dev, poses, args.warmup, args.steps, stream)
        kern16, _, _, acc16 = run_kernels(torch, L, g, frames_dev, poses, args.warmup, args.steps, count=False)
    finally:
        g.close()
    return {"frames_per_s": args.steps / e16, "kernel_ms": kern16,
            "latency_ms": {"p50": float(np.percentile(lat16, 50)), "p99": float(np.percentile(lat16, 99))},
            "accuracy": acc16}


def reference_baseline(args, threads):
    """The reference's own compute shaders on Mesa's llvmpipe
    (oracle/ref_baseline.py), in a child process of its own so that Mesa's
    LLVM never shares this process with the GPU runtime."""
    cmd = [sys.executable, "-m", "oracle.ref_baseline", "--volume", str(args.volume), "--frames",
           str(args.ref_frames), "--warm", str(args.cpu_warm), "--threads", str(threads)]
    p = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        raise RuntimeError("reference baseline exited %d: %s" % (p.returncode, p.stderr[-600:]))
    return json.loads(p.stdout.strip().splitlines()[-1])


def cpu_leg(depth, poses, args):
    """cpu_baseline: the reference's shaders on llvmpipe (kind "reference"),
    with the C oracle's time on the same cores beside it (kind "port"); the
    port alone when the reference library is not built."""
    cfps, threads, cdt = cpu_baseline(depth, poses, args.volume, args.cpu_frames, args.cpu_warm)
    port = {"value": cfps, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": "frames %d-%d of the same %d^3 FP32 sequence (frames 0-%d untimed after the reset), "
                      "C oracle, %d OpenMP threads, %.1f s" % (args.cpu_warm, args.cpu_warm + args.cpu_frames - 1,
                                                                args.volume, args.cpu_warm - 1, threads, cdt)}
    if args.ref_frames <= 0:
        return port
    try:
        r = reference_baseline(args, threads)
    except Exception as e:  # noqa: BLE001 - the port's figure still stands
        port["reference_error"] = "%s: %s" % (type(e).__name__, e)
        return port
    print("bench: reference shaders on llvmpipe %.2f frames/s" % r["frames_per_s"], file=sys.stderr, flush=True)
    return {"value": r["frames_per_s"], "unit": "frames/s", "cores": r["lp_num_threads"], "kind": "reference",
            "sample": "frames %d-%d of the same synthetic sequence at %d^3 RG16I (the reference's volume format): "
                      "the reference's compute shaders (depthToVert, vertToNorm, track, reduce, integrate, raycast) "
                      "on Mesa llvmpipe with %d threads, its host loop (matrices, reduction combine, solve, pose "
                      "update) by the C oracle; frames 0-%d untimed after the reset, %.1f s"
                      % (args.cpu_warm, args.cpu_warm + args.ref_frames - 1, args.volume, r["lp_num_threads"],
                         args.cpu_warm - 1, r["seconds"]),
            "ms_per_frame_by_phase": r["ms_per_frame_by_phase"], "host_cores": r["host_cores"],
            "renderer": r["renderer"], "tracked_frames": r["tracked"], "port": port}


if __name__ == "__main__":
    main()
