"""kf_mesh_generate_noalloc (through the C ABI) against kf_mesh_generate: the
same triangles bit for bit in the caller's device buffer, on two calls in a
row (the second on a smaller volume, reusing the device scratch the first
left behind), a truncated call, and the rejection of a volume beyond the
scratch's size."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
torch = pytest.importorskip("torch")
from oglfusionkinectv2_amd import GFusion, MCubes, kinect_config, synth  # noqa: E402


def _session(volume, frames, storage="f32"):
    depth, gt = synth.sequence(0, frames)
    g = GFusion()
    g.setConfig(kinect_config(volume, 2.0, storage))
    g.allocateBuffers()
    g.Reset(gt[0])
    for f in range(frames):
        g.frame(depth[f].astype(np.float32))
    g.synchronize()
    return g


def _reference_mesh(g):
    mc = MCubes()
    mc.setVolumeTexture(g)
    mc.setIsolevel(0.0)
    n = int(mc.generateMarchingCubes())
    return n, mc.getPosBuffer().reshape(n, 12)


def _noalloc(g, verts_dev, max_tris):
    tri_dev = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    K.check(K.lib().kf_mesh_generate_noalloc(g.session, C.c_float(0.0), C.c_void_p(verts_dev.data_ptr()),
                                             C.c_uint64(max_tris), C.c_void_p(tri_dev.data_ptr())))
    g.synchronize()
    return int(tri_dev.item())


def test_noalloc_mesh_equals_mesh_generate_on_two_calls():
    print("case: kf_mesh_generate_noalloc 128^3 then 64^3", flush=True)
    # one device buffer for both calls, filled with NaN so that writes past
    # the triangle count would show
    verts_dev = torch.full((400000 * 12,), float("nan"), dtype=torch.float32, device="cuda")
    for volume, frames, storage in ((128, 6, "f32"), (64, 4, "f16")):
        g = _session(volume, frames, storage)
        n, ref = _reference_mesh(g)
        assert 1000 < n < 400000, n
        got_n = _noalloc(g, verts_dev, 400000)
        assert got_n == n
        got = verts_dev[: n * 12].cpu().numpy().reshape(n, 12)
        np.testing.assert_array_equal(got.view(np.uint32), ref.reshape(n, 12).view(np.uint32))
        if volume == 64:
            # the rest of the buffer still holds the first call's triangles
            # or the NaN fill, never garbage from this call
            assert np.isnan(verts_dev[399999 * 12:].cpu().numpy()).all()
        g.close()


def test_noalloc_mesh_truncates_to_the_buffer():
    print("case: kf_mesh_generate_noalloc truncated", flush=True)
    g = _session(96, 5)
    n, ref = _reference_mesh(g)
    cap = n // 3
    verts_dev = torch.full(((cap + 8) * 12,), float("nan"), dtype=torch.float32, device="cuda")
    assert _noalloc(g, verts_dev, cap) == n
    out = verts_dev.cpu().numpy()
    np.testing.assert_array_equal(out[: cap * 12].view(np.uint32), ref.reshape(n, 12)[:cap].reshape(-1).view(np.uint32))
    assert np.isnan(out[cap * 12:]).all()
    # a count-only call: no vertex buffer
    tri_dev = torch.zeros((1,), dtype=torch.int64, device="cuda")
    K.check(K.lib().kf_mesh_generate_noalloc(g.session, C.c_float(0.0), None, C.c_uint64(0),
                                             C.c_void_p(tri_dev.data_ptr())))
    g.synchronize()
    assert int(tri_dev.item()) == n
    g.close()


def test_noalloc_mesh_rejects_volumes_beyond_its_scratch():
    g = GFusion()
    cfg = kinect_config(8, 2.0, "f16")
    cfg.volumeSize = (1100, 8, 8)
    g.setConfig(cfg)
    g.allocateBuffers()
    verts_dev = torch.zeros((12,), dtype=torch.float32, device="cuda")
    tri_dev = torch.zeros((1,), dtype=torch.int64, device="cuda")
    rc = K.lib().kf_mesh_generate_noalloc(g.session, C.c_float(0.0), C.c_void_p(verts_dev.data_ptr()),
                                          C.c_uint64(1), C.c_void_p(tri_dev.data_ptr()))
    assert rc == K.KF_ERR_INVALID
    assert b"1024" in K.lib().kf_last_error()
    g.close()
