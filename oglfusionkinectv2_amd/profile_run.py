"""Profiling driver: runs the frame loop on resident synthetic depth for a
fixed number of frames (for rocprofv3 kernel traces and counters), followed by
a plain device copy of known size that calibrates the HBM byte counters.
With --sessions N > 1 the frames are ticks of a saturated batch of N sessions
(kf_batch_frame_u16_dev, every session due every tick, session k on its own
synthetic sequence).

    python -m oglfusionkinectv2_amd.profile_run --volume 512 --storage f32 --frames 40
    python -m oglfusionkinectv2_amd.profile_run --volume 512 --storage f16 --frames 40 --sessions 16
"""
import argparse
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _stream_of(g):
    from oglfusionkinectv2_amd import _lib as K
    h = C.c_void_p()
    K.check(K.lib().kf_get_stream(g.session, C.byref(h)))
    return h.value


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--frames", type=int, default=40)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--stats", type=int, default=0, help="print raycast / integrate statistics every N frames")
    ap.add_argument("--fast-math", type=int, default=1)
    ap.add_argument("--xcd-group", type=int, default=-1, help="raycast XCD cluster (-1: the tuned default)")
    ap.add_argument("--sessions", type=int, default=1, help="> 1: a saturated batch of this many sessions")
    ap.add_argument("--time", type=int, default=0, help="time raycast per frame with events and print the mean")
    ap.add_argument("--wave-trace", type=str, default="",
                    help="comma-separated frames after which one more raycast is traced per wavefront and reported")
    ap.add_argument("--mip", type=int, default=0, help="after the frames, run intensityProjection this many times")
    ap.add_argument("--mesh", type=int, default=0, help="after the frames, run the surface export this many times")
    args = ap.parse_args()

    import torch
    from oglfusionkinectv2_amd import GFusion, kinect_config, synth
    from oglfusionkinectv2_amd import _lib as K

    L = K.lib()
    if args.sessions > 1:
        batch_run(torch, args)
        copy_calibration(torch, args.copy_mib)
        return
    depth, poses = synth.sequence(0, args.frames)
    dev = torch.from_numpy(depth.view(np.int16)).to("cuda")
    g = GFusion()
    g.setConfig(kinect_config(args.volume, 2.0, args.storage))
    g.allocateBuffers()
    g.setOption(K.OPT_FAST_MATH, args.fast_math)
    if args.xcd_group > 0:
        g.setOption(K.OPT_RAYCAST_XCD_GROUP, args.xcd_group)
    g.Reset(poses[0])
    fb = depth[0].nbytes
    traced = set(int(f) for f in args.wave_trace.split(",") if f.strip())
    for i in range(args.frames):
        K.check(L.kf_frame_u16_dev(g.session, C.c_void_p(dev.data_ptr() + i * fb)))
        if i in traced:
            wave_trace_report(g, i)
        if args.stats and i % args.stats == args.stats - 1:
            st = g.raycastStats()
            vr, vw = g.countIntegrateTraffic()
            err = np.linalg.norm(g.getPose()[:3, 3] - poses[i][:3, 3])
            print("frame %d: raycast eval %.1f skip %.1f cells %.1f per ray (max eval %d, max eval+cells %d), "
                  "hits %.3f; integrate read %d written %d; err %.3f m"
                  % (i, st["evaluated"] / 217088, st["skipped"] / 217088, st["bricks"] / 217088, st["max_eval"],
                     st["max_iter"], st["hits"] / 217088, vr, vw, err), flush=True)
    g.synchronize()
    print("frames done:", args.frames, "status", g.status(), flush=True)
    for _ in range(args.mip):
        K.check(L.kf_intensity_projection(g.session))
    if args.mip:
        g.synchronize()
        print("intensity projections done:", args.mip, flush=True)
    tri = C.c_uint64()
    for _ in range(args.mesh):
        K.check(L.kf_mesh_generate(g.session, C.c_float(0.0), C.byref(tri)))
    if args.mesh:
        g.synchronize()
        print("surface exports done:", args.mesh, "triangles", tri.value, flush=True)
    if args.time:
        for grp in (1, 2, 4, 8, 16, 27):
            g.setOption(K.OPT_RAYCAST_XCD_GROUP, grp)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            K.check(L.kf_raycast(g.session))
            g.synchronize()
            stream = torch.cuda.ExternalStream(_stream_of(g))
            e0.record(stream)
            for _ in range(args.time):
                K.check(L.kf_raycast(g.session))
            e1.record(stream)
            torch.cuda.synchronize()
            g.synchronize()
            print("raycast xcd_group %d: %.1f us" % (grp, e0.elapsed_time(e1) * 1000 / args.time), flush=True)
        g.setOption(K.OPT_RAYCAST_XCD_GROUP, args.xcd_group if args.xcd_group > 0 else 4)
        pose = g.getPose()
        for rpb in (1, 2, 4, 6, 8, 14, 27):
            g.setOption(K.OPT_ICP_ROWS_PER_BLOCK, rpb)
            g.setPose(pose)
            K.check(L.kf_track_frame(g.session))
            g.synchronize()
            stream = torch.cuda.ExternalStream(_stream_of(g))
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(args.time):
                g.setPose(pose)
                K.check(L.kf_track_frame(g.session))
            e1.record(stream)
            torch.cuda.synchronize()
            print("icp rows_per_block %d: %.1f us per Track (%d iterations)"
                  % (rpb, e0.elapsed_time(e1) * 1000 / args.time, g.status()["iterations"]), flush=True)
    copy_calibration(torch, args.copy_mib)


def wave_trace_report(g, frame):
    """One more raycast traced per wavefront (kf_raycast_wave_trace): how long
    each wavefront ran against its loop iterations (the largest over its rays
    of fine-step batches of four samples + empty cells entered, as one lane
    counts them), how many wavefronts shared each SIMD, the hybrid schedule's
    four-lane wavefronts apart from the one-lane ones, and for the slowest
    wavefronts the time they ran alone on their SIMD once the others had
    ended."""
    tr = g.raycastWaveTrace()
    iters = tr["iters"].astype(np.float64)
    t0 = (tr["start"] - tr["start"].min()).astype(np.float64) * 0.01  # us (100 MHz)
    t1 = (tr["end"] - tr["start"].min()).astype(np.float64) * 0.01
    dur = t1 - t0
    hw = tr["hw_id"].astype(np.int64)
    simd = ((tr["xcc"].astype(np.int64) * 8 + ((hw >> 13) & 7)) * 2 + ((hw >> 12) & 1)) * 16 + ((hw >> 8) & 15)
    simd = simd * 4 + ((hw >> 4) & 3)
    _, inv, cnt = np.unique(simd, return_inverse=True, return_counts=True)
    print("wave trace frame %d: %d wavefronts on %d SIMDs (wavefronts per SIMD: %s), kernel span %.1f us, "
          "starts within %.1f us" % (frame, len(tr), len(cnt),
                                     {int(a): int(b) for a, b in zip(*np.unique(cnt, return_counts=True))},
                                     t1.max(), t0.max()), flush=True)
    for lanes in (1, 4):
        m = tr["lanes"] == lanes
        if not m.any():
            continue
        A = np.vstack([iters[m], np.ones(m.sum())]).T
        slope, icpt = np.linalg.lstsq(A, dur[m], rcond=None)[0]
        print("  %d lane(s) per ray: %d wavefronts, %d tiles; end p50 %.1f max %.1f us; duration p50 %.1f p99 %.1f "
              "max %.1f us; iterations p50 %d p99 %d max %d; fit duration = %.2f us x iterations %+.1f us"
              % (lanes, m.sum(), len(np.unique(tr["tile"][m])), np.percentile(t1[m], 50), t1[m].max(),
                 np.percentile(dur[m], 50), np.percentile(dur[m], 99), dur[m].max(), np.percentile(iters[m], 50),
                 np.percentile(iters[m], 99), iters[m].max(), slope, icpt), flush=True)
    for k in np.argsort(-t1)[:6]:
        mates = np.nonzero((inv == inv[k]) & (np.arange(len(tr)) != k))[0]
        last = t1[mates].max() if len(mates) else t0[k]
        alone = max(0.0, t1[k] - max(last, t0[k]))
        print("  last to end: tile %d wave %d (%d lane(s) per ray, block %d), %d iterations from %.1f to %.1f us; "
              "%d other wavefronts on its SIMD ended by %.1f us; alone for the last %.1f us"
              % (tr["tile"][k], tr["wave"][k], tr["lanes"][k], tr["block"][k], iters[k], t0[k], t1[k], len(mates),
                 last, alone), flush=True)


def copy_calibration(torch, mib):
    """Five device copies of `mib` MiB: known bytes for the HBM counters."""
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("copy done: %d MiB x 5" % mib, flush=True)


def batch_run(torch, args):
    """`frames` ticks of a saturated batch of `sessions` sessions."""
    from oglfusionkinectv2_amd import GFusion, kinect_config, synth
    from oglfusionkinectv2_amd.batch import SessionBatch
    fb = 512 * 424 * 2
    devs, gs = [], []
    for k in range(args.sessions):
        depth, poses = synth.sequence(k, args.frames)
        devs.append(torch.from_numpy(depth.view(np.int16)).to("cuda"))
        g = GFusion()
        g.setConfig(kinect_config(args.volume, 2.0, args.storage))
        g.allocateBuffers()
        g.Reset(poses[0])
        gs.append(g)
    b = SessionBatch(gs)
    for t in range(args.frames):
        b.frame([d.data_ptr() + t * fb for d in devs])
    b.synchronize()
    print("ticks done:", args.frames, "sessions", args.sessions, "tracked",
          sum(g.status()["tracked"] for g in gs), flush=True)
    b.close()
    for g in gs:
        g.close()


if __name__ == "__main__":
    main()
