"""KF_OPT_INTEGRATE_COLUMNS: the fast integrate's per-column set-up computed
once per frame by k_integrate_columns (one thread per column) gives the same
bits as every z segment's thread computing it: pose, TSDF, reference maps and
status after the same frames, at 128^3 and 512^3, FP32 and FP16."""
import ctypes as C
import hashlib
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
torch = pytest.importorskip("torch")
from oglfusionkinectv2_amd import GFusion, kinect_config, synth  # noqa: E402

F = 10


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


@pytest.mark.parametrize("volume,storage", [(128, "f32"), (512, "f32"), (512, "f16")])
def test_column_setup_once_per_frame_gives_the_same_bits(volume, storage):
    print("case: integrate columns %d^3 %s" % (volume, storage), flush=True)
    depth, poses = synth.sequence(0, F)
    dev = torch.from_numpy(depth.view(np.int16)).to("cuda")
    out = {}
    for cols in (0, 1):
        g = GFusion()
        g.setConfig(kinect_config(volume, 2.0, storage))
        g.allocateBuffers()
        g.setOption(K.OPT_INTEGRATE_COLUMNS, cols)
        assert g.getOption(K.OPT_INTEGRATE_COLUMNS) == cols
        g.Reset(poses[0])
        for f in range(F):
            K.check(K.lib().kf_frame_u16_dev(g.session, C.c_void_p(dev.data_ptr() + f * 512 * 424 * 2)))
        assert g.status()["tracked"] == 1
        out[cols] = _digest(g)
        g.close()
    assert out[0] == out[1]
