"""Autotuner of the frame loop's launch shapes on the GPU it runs on.

The reference dispatches every shader in fixed 32x32 work groups
(gFusion.cpp:514-523, 1268-1269, 1337-1339).  Here the shapes are chosen per
storage type and volume size from measurements; every knob leaves each
result the same bits, so the choice affects time only:

  integrate_zlen       z voxels one integrate thread walks (voxels per
                       thread; 0 = enough segments for about 2M per launch)
  integrate_rows       y rows of an integrate block (64 x rows columns)
  integrate_columns    per-column set-up once per frame in its own launch
                       (1) or in every segment's thread (0)
  raycast_xcd_group    raycast tiles per XCD cluster (blocks whose numbers
                       agree mod 8 run on one XCD and take that many
                       consecutive tiles)
  raycast_lanes        lanes per ray (1: 16x16-ray tiles, 4: 8x8-ray tiles)
  raycast_hybrid       with one lane per ray, the previous launch's heaviest
                       tiles marched four lanes per ray (1) or not (0)
  brick_rebuild_period frames between exact rebuilds of the brick flags

Method: one session per (storage, volume) on synthetic sequence 0; every
candidate runs the same frames from the same reset, and a knob is scored by
the stage it shapes (integrate or raycast kernel duration from the HIP events
in their dispatch packets, mean over the frames after `warm`; the rebuild
period by the GPU time of whole frames over a window of FRAME_WINDOW frames,
a multiple of every candidate period, so that each pays all its rebuilds
whatever the phase of the session's frame counter); coordinate descent in the order
above, each candidate measured `reps` times (best kept), a new value kept
only if it beats the current one by more than `tol`.  Writes the
measurements as JSON and the chosen rows as the C table that
kf_session_create applies (csrc/kf_tuned.inc; rebuild the library after):

    python -m oglfusionkinectv2_amd.autotune --json profiles/autotune_gfx950.json \\
        --inc oglfusionkinectv2_amd/csrc/kf_tuned.inc
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from . import _lib as K  # noqa: E402

KNOBS = [
    ("integrate_zlen", K.OPT_INTEGRATE_ZLEN, "integrate"),
    ("integrate_rows", K.OPT_INTEGRATE_ROWS, "integrate"),
    ("integrate_columns", K.OPT_INTEGRATE_COLUMNS, "integrate"),
    ("raycast_xcd_group", K.OPT_RAYCAST_XCD_GROUP, "raycast"),
    ("raycast_lanes", K.OPT_RAYCAST_LANES, "raycast"),
    ("raycast_hybrid", K.OPT_RAYCAST_HYBRID, "raycast"),
    ("brick_rebuild_period", K.OPT_BRICK_REBUILD_PERIOD, "frame"),
]
DEFAULTS = {"integrate_zlen": 0, "integrate_rows": 4, "integrate_columns": 1, "raycast_xcd_group": 4,
            "raycast_lanes": 1, "raycast_hybrid": 1, "brick_rebuild_period": 16}
STORAGE_CODE = {"f32": 1, "f16": 2}
FRAME_WINDOW = 128  # frames timed for the rebuild period (a multiple of 4, 8, 16, 32 and 64)


def candidates(knob, volume):
    """Values tried for a knob at a volume size (z lengths up to the volume's
    depth)."""
    if knob == "integrate_zlen":
        return [0] + [z for z in (16, 32, 64, 128, 256, 512) if z <= volume]
    if knob == "integrate_rows":
        return [1, 2, 4]
    if knob == "integrate_columns":
        return [0, 1]
    if knob == "raycast_xcd_group":
        return [1, 2, 4, 8, 16]
    if knob == "raycast_lanes":
        return [1, 4]
    if knob == "raycast_hybrid":
        return [0, 1]
    if knob == "brick_rebuild_period":
        return [4, 8, 16, 32, 64]
    raise KeyError(knob)


def choose(times, current, tol=0.01):
    """The value to keep: the fastest, unless it beats `current` by no more
    than the fraction `tol` (measurement noise), in which case `current`."""
    best = min(times, key=times.get)
    if current in times and times[best] >= times[current] * (1.0 - tol):
        return current
    return best


def write_inc(rows, path):
    """The C table of kf_session_create (one initialiser per row)."""
    lines = ["// Launch shapes per storage type and volume size, written by",
             "// oglfusionkinectv2_amd/autotune.py from measurements on MI355X (gfx950).",
             "// {storage, volume, integrate z voxels per thread (0 = automatic), integrate",
             "//  block rows, raycast tiles per XCD cluster, raycast lanes per ray, brick",
             "//  rebuild period, integrate per-column set-up launch, hybrid raycast schedule}"]
    for r in rows:
        c = r["chosen"]
        lines.append("    {%d, %d, %d, %d, %d, %d, %d, %d, %d},  // %s %d^3: integrate %.1f us, raycast %.1f us"
                     % (STORAGE_CODE[r["storage"]], r["volume"], c["integrate_zlen"], c["integrate_rows"],
                        c["raycast_xcd_group"], c["raycast_lanes"], c["brick_rebuild_period"],
                        c["integrate_columns"], c["raycast_hybrid"], r["storage"], r["volume"],
                        1e3 * r["stage_ms"]["integrate"],
                        1e3 * r["stage_ms"]["raycast"]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_inc(path):
    """Rows of a table written by write_inc, as dicts."""
    out = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s.startswith("{"):
                continue
            vals = [int(v) for v in s[1:s.index("}")].split(",")]
            st = {v: k for k, v in STORAGE_CODE.items()}[vals[0]]
            out.append(dict(storage=st, volume=vals[1], integrate_zlen=vals[2], integrate_rows=vals[3],
                            raycast_xcd_group=vals[4], raycast_lanes=vals[5], brick_rebuild_period=vals[6],
                            integrate_columns=vals[7] if len(vals) > 7 else 1,
                            raycast_hybrid=vals[8] if len(vals) > 8 else 1))
    return out


class Runner:
    """One session over `frames` frames of synthetic sequence 0."""

    def __init__(self, torch, volume, storage, frames, warm):
        from .gfusion import GFusion, kinect_config
        from . import synth
        self.torch, self.frames, self.warm = torch, frames, warm
        depth, self.poses = synth.sequence(0, max(frames, warm + FRAME_WINDOW))
        self.dev = torch.from_numpy(depth.view(np.int16)).to("cuda")
        self.stream = torch.cuda.Stream()
        self.g = GFusion(stream=self.stream.cuda_stream)
        self.g.setConfig(kinect_config(volume, 2.0, storage))
        self.g.allocateBuffers()
        self.L = K.lib()

    def _frame(self, i):
        fb = 512 * 424 * 2
        K.check(self.L.kf_frame_u16_dev(self.g.session, C.c_void_p(self.dev.data_ptr() + i * fb)))

    def run(self, opts, metric):
        """ms per frame of `metric` ("integrate", "raycast" or "frame") with
        options `opts`, over frames warm .. frames - 1."""
        g = self.g
        for name, opt, _ in KNOBS:
            g.setOption(opt, opts[name])
        g.setOption(K.OPT_STAGE_TIMING, 1 if metric != "frame" else 0)
        g.Reset(self.poses[0])
        for i in range(self.warm):
            self._frame(i)
        self.torch.cuda.synchronize()
        if metric == "frame":
            e0, e1 = self.torch.cuda.Event(enable_timing=True), self.torch.cuda.Event(enable_timing=True)
            e0.record(self.stream)
            for i in range(self.warm, self.warm + FRAME_WINDOW):
                self._frame(i)
            e1.record(self.stream)
            self.torch.cuda.synchronize()
            return e0.elapsed_time(e1) / FRAME_WINDOW
        acc = []
        for i in range(self.warm, self.frames):
            self._frame(i)
            acc.append(g.stageTimes()[metric])
        g.setOption(K.OPT_STAGE_TIMING, 0)
        return float(np.mean(acc))

    def close(self):
        self.g.close()


def tune(torch, volume, storage, frames=40, warm=15, reps=2, tol=0.01, log=None):
    """Coordinate descent over KNOBS for one (volume, storage); returns the
    row {storage, volume, chosen, measured, stage_ms}."""
    r = Runner(torch, volume, storage, frames, warm)
    cur = dict(DEFAULTS)
    measured = {}
    try:
        for name, _, metric in KNOBS:
            times = {}
            for v in candidates(name, volume):
                o = dict(cur)
                o[name] = v
                times[v] = min(r.run(o, metric) for _ in range(reps))
            measured[name] = {str(k): v for k, v in times.items()}
            cur[name] = choose(times, cur[name], tol)
            if log:
                log("%s %d^3 %s: %s -> %s" % (storage, volume, name,
                                               " ".join("%s:%.4f" % kv for kv in times.items()), cur[name]))
        stage = {m: r.run(cur, m) for m in ("integrate", "raycast")}
        base = {m: r.run(DEFAULTS, m) for m in ("integrate", "raycast")}
    finally:
        r.close()
    return dict(storage=storage, volume=volume, chosen=cur, measured=measured, stage_ms=stage,
                default_stage_ms=base)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--volumes", default="128,256,384,512,768")
    ap.add_argument("--storages", default="f32,f16")
    ap.add_argument("--frames", type=int, default=40)
    ap.add_argument("--warm", type=int, default=15)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--json", default=None, help="write the measurements here")
    ap.add_argument("--inc", default=None, help="write the C table here")
    args = ap.parse_args()
    import torch

    def log(msg):
        print(msg, flush=True)

    rows = []
    t0 = time.perf_counter()
    for st in args.storages.split(","):
        for v in (int(x) for x in args.volumes.split(",")):
            rows.append(tune(torch, v, st, args.frames, args.warm, args.reps, log=log))
    out = {"device": torch.cuda.get_device_name(0), "frames": args.frames, "warm": args.warm, "reps": args.reps,
           "seconds": time.perf_counter() - t0, "rows": rows}
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)
    if args.inc:
        write_inc(rows, args.inc)
    print(json.dumps({"rows": [(r["storage"], r["volume"], r["chosen"]) for r in rows]}))


if __name__ == "__main__":
    main()
