"""Multi-session batches: every session's result is the same alone and in the
mix (bit for bit), with uneven volume sizes and sessions skipping ticks."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
from oglfusionkinectv2_amd import GFusion, kinect_config, synth  # noqa: E402
from oglfusionkinectv2_amd.batch import SessionBatch  # noqa: E402

torch = pytest.importorskip("torch")


def _session(volume, storage="f16", stream=None):
    g = GFusion(stream=stream)
    g.setConfig(kinect_config(volume, 2.0, storage))
    g.allocateBuffers()
    return g


def _state(g):
    v, w = g.getVolume()
    return g.getPose(), v.copy(), w.copy(), g.getVerts(), g.getNorms(), g.status()


@pytest.mark.parametrize("storage,graph,persistent", [("f16", 1, 1), ("f32", 1, 1), ("f16", 0, 0), ("f16", 1, 0),
                                                      ("f16", 0, 1)])
def test_batch_equals_alone(storage, graph, persistent):
    """Each session's bits after 8 frames in a batch (with ticks skipped)
    equal its bits alone, whether the ticks replay captured graphs or launch
    directly and whether Track() runs as one persistent launch for all
    members or one launch per iteration."""
    print("case: batch vs alone %s graph %d persistent %d" % (storage, graph, persistent), flush=True)
    vols = [64, 96, 128]
    F = 8
    seqs = [synth.sequence(i, F) for i in range(3)]
    dev = [torch.from_numpy(d.view(np.int16)).to("cuda") for d, _ in seqs]
    fb = 512 * 424 * 2
    # alone
    alone = []
    for i, n in enumerate(vols):
        g = _session(n, storage)
        g.Reset(seqs[i][1][0])
        for f in range(F):
            K.check(K.lib().kf_frame_u16_dev(g.session, C.c_void_p(dev[i].data_ptr() + f * fb)))
        alone.append(_state(g))
        g.close()
    # in the mix: session 1 skips ticks 2 and 5, session 2 skips tick 0
    gs = [_session(n, storage) for n in vols]
    for i, g in enumerate(gs):
        g.Reset(seqs[i][1][0])
    b = SessionBatch(gs)
    b.setOption(K.OPT_GRAPH, graph)
    b.setOption(K.OPT_ICP_PERSISTENT, persistent)
    nxt = [0, 0, 0]
    tick = 0
    while min(nxt) < F:
        active = [nxt[i] < F for i in range(3)]
        if tick in (2, 5):
            active[1] = False
        if tick == 0:
            active[2] = False
        ptrs = [dev[i].data_ptr() + nxt[i] * fb if active[i] else 0 for i in range(3)]
        b.frame(ptrs, active)
        for i in range(3):
            nxt[i] += 1 if active[i] else 0
        tick += 1
    b.synchronize()
    assert b.getOption(K.OPT_GRAPH) == graph
    assert b.getOption(K.OPT_ICP_PERSISTENT) == persistent
    for i, g in enumerate(gs):
        got = _state(g)
        for a, c in zip(alone[i][:5], got[:5]):
            np.testing.assert_array_equal(a, c)
        assert alone[i][5] == got[5]
    b.close()


def test_batch_rejects_mismatched_sessions():
    a = _session(64, "f16")
    c = _session(64, "f32")
    with pytest.raises(K.KfError):
        SessionBatch([a, c])
    with pytest.raises(K.KfError):
        SessionBatch([a, a])
