"""Host logic of the serving loop (admission against an HBM budget, free-list
reuse, continuous batching) over a recording backend: no GPU."""
import math

import numpy as np

from oglfusionkinectv2_amd import serve


class FakeBackend:
    def __init__(self, bytes_of):
        self.bytes_of = bytes_of
        self.live_handles = set()
        self.created = []
        self.destroyed = []
        self.starts = []
        self.member_calls = []
        self.frames = []  # (members, active) per launch
        self.n = 0
        self.gpu_s = 0.0  # simulated duration of one launch (0: complete at once)
        self.clock = None
        self.finish = []

    def bytes(self, volume):
        return self.bytes_of[volume]

    def create(self, volume):
        self.n += 1
        h = ("h", self.n, volume)
        self.live_handles.add(h)
        self.created.append(h)
        return h

    def destroy(self, h):
        self.live_handles.remove(h)
        self.destroyed.append(h)

    def start(self, h, seq):
        self.starts.append((h, seq))

    def members(self, hs):
        assert 1 <= len(hs) <= serve.MAX_BATCH
        assert len(set(hs)) == len(hs)
        assert all(h in self.live_handles for h in hs)
        self.member_calls.append(list(hs))

    def depth_ptr(self, seq, i):
        return (seq, i)

    def frame(self, ptrs, active):
        self.frames.append((list(self.member_calls[-1]), list(ptrs), list(active)))
        if self.gpu_s:
            start = max(self.clock(), self.finish[-1] if self.finish else 0.0)
            self.finish.append(start + self.gpu_s)

    def mark(self):
        return len(self.frames)

    def done(self, mark):
        # with a simulated GPU (gpu_s > 0) launch `mark` finishes gpu_s after
        # the later of its launch time and the previous launch's end
        if not self.gpu_s:
            return True
        return self.clock() >= self.finish[mark - 1]


class VirtualClock:
    """Time advances only in sleep(); with dt > 0 the loop wakes on a dt grid
    (a host that polls every dt seconds)."""

    def __init__(self, dt):
        self.t, self.dt = 0.0, dt

    def __call__(self):
        return self.t

    def sleep(self, w):
        t = self.t + max(w, 0.0)
        self.t = math.ceil(t / self.dt - 1e-9) * self.dt if self.dt else t


def run_virtual(arrivals, budget, bytes_of, seconds=1e9, dt=0.0, on_retire=None, sequences=4, gpu_s=0.0,
                max_inflight=2):
    be = FakeBackend(bytes_of)
    srv = serve.FusionServer(be, budget, sequences, on_retire=on_retire, max_inflight=max_inflight)
    clk = VirtualClock(dt)
    be.gpu_s, be.clock = gpu_s, clk
    used_trace = []
    orig_tick = srv.tick

    def tick(now):
        r = orig_tick(now)
        used_trace.append(srv.used)
        assert srv.used <= budget
        return r

    srv.tick = tick
    serve.drive(srv, arrivals, seconds, clk, clk.sleep)
    return srv, be, used_trace


def test_every_frame_once_in_order_and_on_time():
    arr = [(0.0, 64, 30.0, 20), (0.01, 128, 15.0, 12), (0.2, 64, 30.0, 15)]
    srv, be, _ = run_virtual(arr, 1000, {64: 10, 128: 40})
    seen = {}
    for (mark, dues) in srv.records:
        for sid, d in dues:
            seen.setdefault(sid, []).append(d)
    for sid, (t, v, r, f) in enumerate(arr):
        d = seen[sid]
        assert len(d) == f
        assert np.allclose(np.diff(d), 1.0 / r)  # frames in order, one per period
    # every launch carries exactly the sessions with a frame due: the frame
    # index each active member gets is the next one of its sequence
    nxt = {}
    for members, ptrs, active in be.frames:
        assert len(members) == len(ptrs) == len(active)
        for h, p, a in zip(members, ptrs, active):
            if a:
                seq, i = p
                assert i == nxt.get(h, 0)
                nxt[h] = i + 1


def test_continuous_batching_joins_next_launch():
    # two sessions with the same period, phases 1 ms apart: a loop that wakes
    # every 10 ms finds both frames due and sends them in one launch
    arr = [(0.002, 64, 10.0, 5), (0.003, 64, 10.0, 5)]
    srv, be, _ = run_virtual(arr, 100, {64: 10}, dt=0.01)
    sizes = [len(d) for _, d in srv.records]
    assert sizes == [2] * 5
    # waking exactly on each due time sends them one by one
    srv, be, _ = run_virtual(arr, 100, {64: 10}, dt=0.0)
    assert [len(d) for _, d in srv.records] == [1] * 10


def test_budget_blocks_then_admits_fifo():
    # budget fits two 128^3 sessions; the third waits for the first to leave
    arr = [(0.0, 128, 30.0, 10), (0.0, 128, 30.0, 40), (0.05, 128, 30.0, 5), (0.06, 64, 30.0, 5)]
    srv, be, trace = run_virtual(arr, 80, {64: 10, 128: 40})
    assert max(trace) <= 80
    assert srv.admitted == 4
    assert srv.waits > 0
    # FIFO: session 3 (small, would fit early) is not admitted before session 2
    first_frame = {}
    for k, (_, dues) in enumerate(srv.records):
        for sid, _d in dues:
            first_frame.setdefault(sid, k)
    assert first_frame[2] <= first_frame[3]
    # the third 128^3 session reused the first one's buffers
    assert srv.allocated == 3 and srv.reused == 1


def test_free_list_released_for_other_sizes():
    arr = [(0.0, 128, 30.0, 5), (1.0, 256, 30.0, 5)]
    srv, be, trace = run_virtual(arr, 100, {128: 40, 256: 90})
    assert srv.admitted == 2
    assert len(be.destroyed) == 1 and be.destroyed[0][2] == 128
    assert max(trace) <= 100


def test_batch_limit():
    arr = [(0.0, 64, 30.0, 3) for _ in range(serve.MAX_BATCH + 6)]
    srv, be, _ = run_virtual(arr, 10 ** 9, {64: 1})
    assert srv.max_live == serve.MAX_BATCH
    assert srv.admitted == serve.MAX_BATCH + 6
    assert max(len(m) for m in be.member_calls) == serve.MAX_BATCH


def test_on_retire_sees_every_session_once():
    got = []
    arr = serve.poisson_sessions(seed=3, total=12, mean_gap_s=0.05, sizes=(64, 96), rates=(30.0,),
                                 frames_range=(3, 9))
    srv, be, _ = run_virtual(arr, 10 ** 9, {64: 1, 96: 2}, on_retire=lambda s: got.append((s.sid, s.done)))
    assert sorted(s for s, _ in got) == list(range(12))
    assert all(done == arr[sid][3] for sid, done in got)


def test_poisson_generator_matches_config():
    arr = serve.poisson_sessions()
    assert len(arr) == 64
    t = np.array([a[0] for a in arr])
    assert np.all(np.diff(t) > 0)
    assert 0.5 < np.mean(np.diff(t)) < 2.0  # mean gap 1 s
    assert {a[1] for a in arr} <= {256, 384, 512, 768}
    assert {a[2] for a in arr} <= {15.0, 30.0}
    assert all(200 <= a[3] <= 900 for a in arr)


def test_busy_gpu_batches_the_frames_that_fall_due():
    """Sixteen 30 Hz sessions with first frames spread over one period and a
    GPU that needs 10 ms per launch: with one launch in flight at a time the
    frames that fall due meanwhile ride the next launch together (several per
    launch, every frame once and in order); with no cap every frame would be
    its own launch queued behind the others."""
    arr = [(k / (16 * 30.0), 64, 30.0, 30) for k in range(16)]
    for cap in (1, 2):
        srv, be, _ = run_virtual(arr, 10 ** 9, {64: 1}, gpu_s=0.010, max_inflight=cap)
        sizes = [len(d) for _, d in srv.records]
        assert sum(sizes) == 16 * 30
        assert srv.blocked > 0
        assert np.mean(sizes) >= 3.0, (cap, np.mean(sizes))
        # a launch never holds more work than max_inflight launches behind it
        # can make wait: no frame is launched more than (cap + 1) launches
        # late, i.e. its launch ends within (cap + 1) * gpu_s + one period
        for k, (m, dues) in enumerate(srv.records):
            end = be.finish[m - 1]
            for sid, d in dues:
                assert end - d <= (cap + 1) * 0.010 + 1.0 / 30.0 + 1e-9
        seen = {}
        for _, dues in srv.records:
            for sid, d in dues:
                seen.setdefault(sid, []).append(d)
        for sid in range(16):
            assert np.allclose(np.diff(seen[sid]), 1.0 / 30.0)
    srv, be, _ = run_virtual(arr, 10 ** 9, {64: 1}, gpu_s=0.010, max_inflight=0)
    assert max(len(d) for _, d in srv.records) == 1
