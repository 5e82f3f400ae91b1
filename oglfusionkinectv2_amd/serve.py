"""Serving loop: many concurrent reconstruction sessions on one GPU.

Sessions arrive over time (a Poisson process, or all at t = 0), are admitted
against an HBM budget (kf_session_bytes), stream 512x424 depth frames at their
own rate and leave after their frame count.  Each loop pass the server
gathers every session with a frame due and runs them as ONE batched frame
(kf_batch_frame_u16_dev): continuous batching, so frames that fall due while
earlier ones are still on the GPU join the next launch instead of waiting for
a fixed batch to fill.  Nothing on the request path blocks on the GPU: batch
membership changes are stream-ordered, and a leaving session's buffers go to a
free list and are reused (kf_reset) by the next admission of the same volume
size, so only a session of a new size allocates.

Per-frame latency runs from the frame's due time to the GPU completion of the
launch that carried it (a HIP event on the batch stream, read after the run),
reported as p50 / p99 beside the aggregate frames/s.

The host logic (admission, free list, which sessions ride each launch) lives
in `FusionServer`, over a backend with the device calls; `GpuBackend` is the
kfusion one.

    python -m oglfusionkinectv2_amd.serve --sessions 16 --volume 512 --rate 30 --seconds 10
    python -m oglfusionkinectv2_amd.serve --poisson --seconds 60
"""
import argparse
import ctypes as C
import heapq
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from . import _lib as K  # noqa: E402

MAX_BATCH = 64  # kf_batch_create limit


# ---------------------------------------------------------------------------
# host logic
# ---------------------------------------------------------------------------
class Session:
    """One admitted session: `frames` depth frames of sequence `seq` at `rate`
    Hz, the first due at `t_start`."""

    def __init__(self, sid, volume, rate, frames, t_start, seq, handle, nbytes):
        self.sid, self.volume, self.rate, self.frames = sid, volume, rate, frames
        self.t_start, self.seq, self.handle, self.bytes = t_start, seq, handle, nbytes
        self.done = 0  # frames submitted

    def due(self, k):
        return self.t_start + k / self.rate

    def next_due(self):
        return self.due(self.done) if self.done < self.frames else None


class FusionServer:
    """Admission, free list and continuous batching over a backend with
    bytes(volume), create(volume), destroy(h), start(h, seq), members(hs),
    frame(ptrs, active), depth_ptr(seq, i), mark() and done(mark).

    Admission is first-come first-served: a request waits while the live and
    free-listed sessions plus its own bytes would exceed the budget (free-list
    entries of other sizes are released first) or the batch is full.

    At most `max_inflight` launches are queued on the GPU at once (0: no
    limit): while that many are still running, frames that fall due wait on
    the host and all ride the next launch together, so under load the batch
    grows instead of the stream filling with one-frame launches."""

    def __init__(self, backend, budget_bytes, sequences, on_retire=None, max_inflight=2):
        self.b = backend
        self.budget = budget_bytes
        self.sequences = sequences
        self.on_retire = on_retire
        self.max_inflight = max_inflight
        self.inflight = []  # marks of launches not yet seen complete, oldest first
        self.blocked = 0  # passes that found every launch slot busy
        self.live = []
        self.free = {}  # volume -> [handle]
        self.waiting = []  # FIFO of (sid, volume, rate, frames, t_request)
        self.used = 0  # bytes of live + free-listed sessions
        self.records = []  # (mark, [(sid, due)])
        self.admitted = 0
        self.allocated = 0
        self.reused = 0
        self.max_live = 0
        self.max_used = 0
        self.waits = 0

    def _bytes(self, volume):
        return self.b.bytes(volume)

    def request(self, sid, volume, rate, frames, now):
        self.waiting.append((sid, volume, rate, frames, now))
        self._admit(now)

    def _take(self, volume, nbytes):
        lst = self.free.get(volume)
        if lst:
            self.reused += 1
            return lst.pop()
        # release free-listed sessions of other sizes until the new one fits
        for v in sorted(self.free, key=lambda v: -self._bytes(v)):
            while self.free[v] and self.used + nbytes > self.budget:
                self.b.destroy(self.free[v].pop())
                self.used -= self._bytes(v)
        if self.used + nbytes > self.budget:
            return None
        h = self.b.create(volume)
        self.used += nbytes
        self.allocated += 1
        return h

    def _admit(self, now):
        changed = False
        while self.waiting and len(self.live) < MAX_BATCH:
            sid, volume, rate, frames, t_req = self.waiting[0]
            nbytes = self._bytes(volume)
            h = self._take(volume, nbytes)
            if h is None:
                self.waits += 1
                break
            self.waiting.pop(0)
            seq = sid % self.sequences
            self.b.start(h, seq)
            self.live.append(Session(sid, volume, rate, frames, max(now, t_req), seq, h, nbytes))
            self.admitted += 1
            changed = True
        if changed:
            self.b.members([s.handle for s in self.live])
            self.max_live = max(self.max_live, len(self.live))
            self.max_used = max(self.max_used, self.used)

    def _retire(self, now):
        gone = [s for s in self.live if s.done >= s.frames]
        if not gone:
            return
        self.live = [s for s in self.live if s.done < s.frames]
        if self.live:
            self.b.members([s.handle for s in self.live])
        for s in gone:
            if self.on_retire:
                self.on_retire(s)
            self.free.setdefault(s.volume, []).append(s.handle)
        self._admit(now)

    def next_due(self):
        return min((d for d in (s.next_due() for s in self.live) if d is not None), default=None)

    def busy(self):
        """True while max_inflight launches are still running."""
        if not self.max_inflight:
            return False
        while self.inflight and self.b.done(self.inflight[0]):
            self.inflight.pop(0)
        return len(self.inflight) >= self.max_inflight

    def tick(self, now):
        """One batched frame for every session with a frame due by `now`
        (none while every launch slot is busy); returns the number of frames
        launched."""
        if self.busy():
            self.blocked += 1
            return 0
        active, ptrs, dues = [], [], []
        for s in self.live:
            d = s.next_due()
            if d is not None and d <= now:
                active.append(True)
                ptrs.append(self.b.depth_ptr(s.seq, s.done))
                dues.append((s.sid, d))
                s.done += 1
            else:
                active.append(False)
                ptrs.append(0)
        if not dues:
            return 0
        self.b.frame(ptrs, active)
        m = self.b.mark()
        self.records.append((m, dues))
        if self.max_inflight:
            self.inflight.append(m)
        self._retire(now)
        return len(dues)

    def close(self):
        for s in self.live:
            self.free.setdefault(s.volume, []).append(s.handle)
        self.live = []
        for v, lst in self.free.items():
            for h in lst:
                self.b.destroy(h)
        self.free = {}


def drive(server, arrivals, seconds, clock, sleep):
    """Runs the request loop until `seconds` of `clock()` have passed or every
    session has finished.  arrivals: (t, volume, rate_hz, frames) per session
    (session id = position)."""
    q = [(t, i, v, r, f) for i, (t, v, r, f) in enumerate(arrivals)]
    heapq.heapify(q)
    launched = 0
    while True:
        now = clock()
        if now >= seconds:
            break
        while q and q[0][0] <= now:
            t, i, v, r, f = heapq.heappop(q)
            server.request(i, v, r, f, t)
        launched += server.tick(now)
        nd = server.next_due()
        cands = [x for x in (nd, q[0][0] if q else None) if x is not None]
        if not cands:  # nothing live or arriving: a waiting request can never fit
            break
        wait = min(cands) - clock()
        if wait > 0:
            sleep(wait)
        elif server.busy():
            sleep(1e-4)  # frames are due but every launch slot is busy: poll
    return launched


# ---------------------------------------------------------------------------
# kfusion backend
# ---------------------------------------------------------------------------
def session_bytes(volume, storage):
    from .gfusion import kinect_config, to_kf_config
    kc = to_kf_config(kinect_config(volume, 2.0, storage))
    n = C.c_size_t()
    K.check(K.lib().kf_session_bytes(C.byref(kc), C.byref(n)))
    return n.value


class DepthPool:
    """Device-resident synthetic sequences: sequence k holds `frames` frames of
    synthetic session k.  Frame i of a session plays its sequence forward then
    backward (ping-pong) so that the camera path stays smooth however long the
    session lasts."""

    def __init__(self, torch, sequences, frames):
        from . import synth
        self.frames = frames
        self.poses, self.dev = [], []
        for k in range(sequences):
            d, p = synth.sequence(k, frames)
            self.dev.append(torch.from_numpy(d.view(np.int16)).to("cuda"))
            self.poses.append(p)
        self.frame_bytes = 512 * 424 * 2

    def index(self, i):
        L = self.frames
        if L == 1:
            return 0
        j = i % (2 * L - 2)
        return j if j < L else 2 * L - 2 - j

    def ptr(self, seq, i):
        return self.dev[seq].data_ptr() + self.index(i) * self.frame_bytes


class GpuBackend:
    """Sessions on their own streams (allocation and reset never wait for the
    batch), one kf_batch on the serving stream."""

    def __init__(self, torch, pool, storage="f16", options=None, batch_options=None):
        self.torch, self.pool, self.storage = torch, pool, storage
        self.options = dict(options or {})
        self.batch_options = dict(batch_options or {})
        self.stream = torch.cuda.Stream()
        self.batch = None
        self._bytes = {}

    def bytes(self, volume):
        if volume not in self._bytes:
            self._bytes[volume] = session_bytes(volume, self.storage)
        return self._bytes[volume]

    def create(self, volume):
        from .gfusion import GFusion, kinect_config
        g = GFusion()
        g.setConfig(kinect_config(volume, 2.0, self.storage))
        g.allocateBuffers()
        for k, v in self.options.items():
            g.setOption(k, v)
        return g

    def destroy(self, g):
        g.close()

    def start(self, g, seq):
        # a reused session may still be read by frames queued on the batch
        # stream: its own stream waits for them before the reset
        h = C.c_void_p()
        K.check(K.lib().kf_get_stream(g.session, C.byref(h)))
        own = self.torch.cuda.ExternalStream(h.value)
        own.wait_stream(self.stream)
        g.Reset(self.pool.poses[seq][0])

    def members(self, gs):
        from .batch import SessionBatch
        if self.batch is None:
            self.batch = SessionBatch(gs, self.stream.cuda_stream)
            for k, v in self.batch_options.items():
                self.batch.setOption(k, v)
        else:
            self.batch.set_sessions(gs)

    def depth_ptr(self, seq, i):
        return self.pool.ptr(seq, i)

    def frame(self, ptrs, active):
        self.batch.frame(ptrs, active)

    def mark(self):
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        return ev

    def done(self, ev):
        return ev.query()

    def close(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None


def latency_stats(records, origin, warm_s=0.5):
    """records: (end event, [(sid, due)]) -> per-frame latency in ms for
    frames due after `warm_s`; origin: event recorded at clock time 0."""
    lat = []
    for ev, dues in records:
        done = origin.elapsed_time(ev) / 1000.0
        lat.extend((done - d) * 1000.0 for _, d in dues if d >= warm_s)
    return np.asarray(lat, dtype=np.float64)


def serve(torch, arrivals, seconds, budget_bytes=200e9, storage="f16", sequences=8, pool_frames=64,
          warm_s=0.5, options=None, max_inflight=2, batch_options=None):
    """Real-time run of `arrivals` for at most `seconds`.  Returns a dict of
    frames/s, latency percentiles and admission counts."""
    pool = DepthPool(torch, sequences, pool_frames)
    be = GpuBackend(torch, pool, storage, options, batch_options)
    srv = FusionServer(be, budget_bytes, sequences, max_inflight=max_inflight)
    for v in sorted({a[1] for a in arrivals}):  # warm the per-size code paths
        g = be.create(v)
        be.destroy(g)
    torch.cuda.synchronize()
    origin = torch.cuda.Event(enable_timing=True)
    origin.record(be.stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0  # noqa: E731

    def sleep(w):
        time.sleep(min(w, 0.002) if w > 0.0002 else 0)

    drive(srv, arrivals, seconds, clock, sleep)
    torch.cuda.synchronize()
    span = clock()
    lat = latency_stats(srv.records, origin, warm_s)
    frames = sum(len(d) for _, d in srv.records)
    launches = len(srv.records)
    srv.close()
    be.close()
    if lat.size == 0:
        lat = np.array([float("nan")])
    return {
        "frames": int(frames),
        "seconds": span,
        "frames_per_s": frames / span if span > 0 else 0.0,
        "launches": launches,
        "frames_per_launch": frames / launches if launches else 0.0,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(np.max(lat))},
        "sessions_requested": len(arrivals),
        "sessions_admitted": srv.admitted,
        "sessions_allocated": srv.allocated,
        "sessions_reused": srv.reused,
        "max_concurrent": srv.max_live,
        "max_hbm_GB": srv.max_used / 1e9,
        "hbm_budget_GB": budget_bytes / 1e9,
        "storage": storage,
        "max_inflight": max_inflight,
        "passes_blocked": srv.blocked,
    }


def saturate(torch, volumes, ticks, warmup=10, storage="f16", sequences=8, pool_frames=64, options=None,
             batch_options=None):
    """Every session has a frame due every tick (offline throughput): one
    batched frame per tick for `ticks` ticks.  Returns aggregate frames/s and
    voxel updates/s."""
    pool = DepthPool(torch, sequences, pool_frames)
    be = GpuBackend(torch, pool, storage, options, batch_options)
    hs = [be.create(v) for v in volumes]
    for i, g in enumerate(hs):
        be.start(g, i % sequences)
    be.members(hs)
    act = [True] * len(hs)
    for t in range(warmup):
        be.frame([pool.ptr(i % sequences, t) for i in range(len(hs))], act)
    torch.cuda.synchronize()
    e0, e1 = be.mark(), None
    for t in range(warmup, warmup + ticks):
        be.frame([pool.ptr(i % sequences, t) for i in range(len(hs))], act)
    e1 = be.mark()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    tracked = [g.status()["tracked"] for g in hs]
    for g in hs:
        be.destroy(g)
    be.close()
    n = len(volumes)
    fps = n * ticks / (ms / 1000.0)
    return {"sessions": n, "volumes": list(volumes), "ticks": ticks, "ms_per_tick": ms / ticks,
            "frames_per_s": fps, "voxel_updates_per_s": sum(v ** 3 for v in volumes) * ticks / (ms / 1000.0),
            "tracked_last": int(sum(tracked)), "storage": storage}


def fixed_sessions(n, volume, rate, seconds):
    """n sessions that all stay for the whole run, first frames spread over one
    frame period."""
    return [(k / (n * rate), volume, rate, int(rate * seconds) + 10) for k in range(n)]


def poisson_sessions(seed=7, total=64, mean_gap_s=1.0, sizes=(256, 384, 512, 768), rates=(15.0, 30.0),
                     frames_range=(200, 901)):
    """Poisson arrivals (exponential gaps), size and rate drawn uniformly, frame
    count uniform in frames_range."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for _ in range(total):
        t += rng.exponential(mean_gap_s)
        out.append((float(t), int(rng.choice(sizes)), float(rng.choice(rates)), int(rng.integers(*frames_range))))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sessions", type=int, default=16)
    ap.add_argument("--volume", type=int, default=512)
    ap.add_argument("--rate", type=float, default=30.0)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--storage", default="f16")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--poisson", action="store_true", help="64 Poisson arrivals at 1/s, mixed sizes and rates")
    ap.add_argument("--saturate", type=int, default=0, help="offline: N ticks with every session due")
    args = ap.parse_args()
    import torch
    if args.saturate:
        out = saturate(torch, [args.volume] * args.sessions, args.saturate, storage=args.storage)
    else:
        arr = (poisson_sessions() if args.poisson
               else fixed_sessions(args.sessions, args.volume, args.rate, args.seconds))
        out = serve(torch, arr, args.seconds, args.budget_gb * 1e9, args.storage)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
