"""Every value the autotuner may pick leaves the frame loop's results the
same bits (pose, TSDF, reference maps, status) as the defaults, alone and
for a batch whose members were tuned differently; and a short tuning run
completes and picks values among its candidates."""
import ctypes as C
import hashlib
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
torch = pytest.importorskip("torch")
from oglfusionkinectv2_amd import GFusion, autotune, kinect_config, synth  # noqa: E402
from oglfusionkinectv2_amd.batch import SessionBatch  # noqa: E402

F = 8


def _digest(g):
    h = hashlib.blake2b(digest_size=16)
    h.update(g.getPose().tobytes())
    v, w = g.getVolume()
    h.update(v.tobytes())
    h.update(w.tobytes())
    h.update(g.getVerts().tobytes())
    h.update(g.getNorms().tobytes())
    h.update(json.dumps(g.status(), sort_keys=True).encode())
    return h.hexdigest()


def _run(volume, storage, opts, dev, poses):
    g = GFusion()
    g.setConfig(kinect_config(volume, 2.0, storage))
    g.allocateBuffers()
    for name, opt, _ in autotune.KNOBS:
        g.setOption(opt, opts[name])
    g.Reset(poses[0])
    for f in range(F):
        K.check(K.lib().kf_frame_u16_dev(g.session, C.c_void_p(dev.data_ptr() + f * 512 * 424 * 2)))
    d = _digest(g)
    g.close()
    return d


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_every_knob_value_gives_the_same_bits(storage):
    print("case: autotune knobs 128^3 %s" % storage, flush=True)
    depth, poses = synth.sequence(0, F)
    dev = torch.from_numpy(depth.view(np.int16)).to("cuda")
    ref = _run(128, storage, autotune.DEFAULTS, dev, poses)
    for name, _, _ in autotune.KNOBS:
        for v in autotune.candidates(name, 128):
            if v == autotune.DEFAULTS[name]:
                continue
            o = dict(autotune.DEFAULTS)
            o[name] = v
            assert _run(128, storage, o, dev, poses) == ref, (name, v)


def test_batch_of_differently_tuned_members_equals_alone():
    print("case: batch of differently tuned members", flush=True)
    vols = [64, 128]
    seqs = [synth.sequence(i, F) for i in range(2)]
    dev = [torch.from_numpy(d.view(np.int16)).to("cuda") for d, _ in seqs]
    tuned = [dict(autotune.DEFAULTS, integrate_rows=1, raycast_xcd_group=2, integrate_zlen=16),
             dict(autotune.DEFAULTS, raycast_lanes=4, raycast_xcd_group=8, brick_rebuild_period=8)]
    alone = [_run(vols[i], "f16", tuned[i], dev[i], seqs[i][1]) for i in range(2)]
    gs = []
    for i in range(2):
        g = GFusion()
        g.setConfig(kinect_config(vols[i], 2.0, "f16"))
        g.allocateBuffers()
        for name, opt, _ in autotune.KNOBS:
            g.setOption(opt, tuned[i][name])
        g.Reset(seqs[i][1][0])
        gs.append(g)
    b = SessionBatch(gs)
    for f in range(F):
        b.frame([dev[i].data_ptr() + f * 512 * 424 * 2 for i in range(2)])
    b.synchronize()
    for i in range(2):
        assert _digest(gs[i]) == alone[i], i
    b.close()
    for g in gs:
        g.close()


def test_short_tuning_run():
    print("case: tune 64^3 f16", flush=True)
    r = autotune.tune(torch, 64, "f16", frames=10, warm=5, reps=1)
    for k in autotune.DEFAULTS:
        assert r["chosen"][k] in autotune.candidates(k, 64)
    assert r["stage_ms"]["integrate"] > 0 and r["stage_ms"]["raycast"] > 0
