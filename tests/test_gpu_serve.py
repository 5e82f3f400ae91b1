"""Serving loop on the GPU: sessions of different volume sizes join and leave a
continuous batch (buffers reused through the free list); every session's
final pose and TSDF equal those of the same frames run alone."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class StepClock:
    """Virtual clock: the loop wakes every `dt` seconds, so frames of several
    sessions fall due together and share launches."""

    def __init__(self, dt):
        self.t, self.dt = 0.0, dt

    def __call__(self):
        return self.t

    def sleep(self, w):
        self.t += self.dt


def test_served_sessions_equal_alone():
    import torch

    from oglfusionkinectv2_amd import GFusion, kinect_config, serve
    from oglfusionkinectv2_amd import _lib as K

    print("serve parity: volumes 64/96/128 f16, 7 sessions", flush=True)
    pool = serve.DepthPool(torch, 3, 12)
    be = serve.GpuBackend(torch, pool, "f16")
    results = {}

    def on_retire(s):
        torch.cuda.synchronize()
        v, w = s.handle.getVolume()
        results[s.sid] = (s.volume, s.seq, s.done, s.handle.getPose(), v.copy(), w.copy())

    # 7 sessions, at most 3 live (budget), sizes repeat so buffers are reused
    arr = [(0.00, 64, 30.0, 14), (0.01, 96, 30.0, 9), (0.02, 128, 15.0, 8), (0.05, 64, 30.0, 10),
           (0.10, 96, 15.0, 7), (0.12, 128, 30.0, 12), (0.15, 64, 15.0, 6)]
    b = {v: be.bytes(v) for v in (64, 96, 128)}
    srv = serve.FusionServer(be, b[64] + b[96] + b[128], 3, on_retire=on_retire)
    clk = StepClock(0.02)
    serve.drive(srv, arr, 100.0, clk, clk.sleep)
    torch.cuda.synchronize()
    assert srv.reused > 0 and max(len(d) for _, d in srv.records) > 1
    srv.close()
    be.close()
    assert sorted(results) == list(range(len(arr)))

    L = K.lib()
    for sid, (vol, seq, n, pose, v, w) in sorted(results.items()):
        g = GFusion()
        g.setConfig(kinect_config(vol, 2.0, "f16"))
        g.allocateBuffers()
        g.Reset(pool.poses[seq][0])
        for i in range(n):
            K.check(L.kf_frame_u16_dev(g.session, C.c_void_p(pool.ptr(seq, i))))
        P = g.getPose()
        va, wa = g.getVolume()
        g.close()
        assert np.array_equal(P, pose), "session %d (%d^3): pose differs" % (sid, vol)
        assert np.array_equal(va.view(np.uint16), v.view(np.uint16)), "session %d: TSDF differs" % sid
        assert np.array_equal(wa.view(np.uint16), w.view(np.uint16)), "session %d: weight differs" % sid


def test_saturated_batch_runs():
    import torch

    from oglfusionkinectv2_amd import serve

    print("serve saturate: 3 sessions 64/96/128 f16, 5 ticks", flush=True)
    r = serve.saturate(torch, [64, 96, 128], ticks=5, warmup=2, sequences=3, pool_frames=8)
    assert r["frames_per_s"] > 0
    assert r["tracked_last"] == 3
