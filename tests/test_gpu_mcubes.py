"""Surface export on the GPU (kf_mesh_*, MCubes) against the oracle's
histopyramid marching cubes (histoPyramids.cs / traverseHistoPyramids.cs):
same triangles in the same order, bit for bit (both evaluate the same FP32
vertex expressions; the GPU skips bricks whose dilated occupancy flag is clear,
which must not change anything), and the STL / PLY files of exportMesh /
exportPointCloud."""
import struct

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

K = pytest.importorskip("oglfusionkinectv2_amd._lib")
from oglfusionkinectv2_amd import GFusion, MCubes, MCubeConfig, kinect_config  # noqa: E402

STORAGE_ORACLE = {"rg16i": oracle.RG16I, "f32": oracle.F32X2, "f16": oracle.F16X2}


def planes(vol, storage):
    if storage == "f16":
        return vol[..., 0].view(np.float16), vol[..., 1].view(np.float16)
    return vol[..., 0], vol[..., 1]


def gpu_mesh(g, iso=0.0):
    mc = MCubes()
    cfg = MCubeConfig()
    cfg.gridSize = g.configuration.volumeSize
    mc.setConfig(cfg)
    mc.setVolumeTexture(g)
    mc.init()
    mc.setIsolevel(iso)
    mc.generateMarchingCubes()
    return mc, mc.getPosBuffer()


@pytest.mark.parametrize("storage", ["rg16i", "f32", "f16"])
def test_mesh_of_a_fused_volume_matches_oracle(o32, seq20, storage):
    print("case: mesh 128^3 %s" % storage, flush=True)
    depth, gt = seq20
    cfg = kinect_config(128, 2.0, storage)
    s = oracle.OracleSession(o32, 128, 2.0, cfg.mu, STORAGE_ORACLE[storage])
    s.reset(gt[0])
    for f in range(6):
        s.frame(depth[f].astype(np.float32))
    vol = s.volume()
    g = GFusion()
    g.setConfig(cfg)
    g.allocateBuffers()
    g.Reset(gt[0])
    g.setVolume(*planes(vol, storage))
    mc, V = gpu_mesh(g)
    W = o32.mc_extract(vol, STORAGE_ORACLE[storage])
    assert len(W) > 1000
    assert mc.getNumberTriangles() == 3 * len(W)
    np.testing.assert_array_equal(V, W)


def test_mesh_of_an_analytic_sphere_256(o32):
    print("case: mesh sphere 256^3 f32", flush=True)
    n = 256
    z, y, x = np.meshgrid(np.arange(n, dtype=np.float32), np.arange(n, dtype=np.float32),
                          np.arange(n, dtype=np.float32), indexing="ij")
    r = np.sqrt((x - 127.3) ** 2 + (y - 121.6) ** 2 + (z - 130.2) ** 2)
    vol = np.zeros((n, n, n, 2), np.float32)
    vol[..., 0] = np.clip((r - 60.0) * 0.02, -0.1, 0.2)
    vol[..., 1] = 1
    del x, y, z, r
    g = GFusion()
    g.setConfig(kinect_config(n, 2.0, "f32"))
    g.allocateBuffers()
    g.setVolume(*planes(vol, "f32"))
    _, V = gpu_mesh(g)
    W = o32.mc_extract(vol, oracle.F32X2)
    np.testing.assert_array_equal(V, W)
    rad = np.linalg.norm(V[:, :, :3] - np.float32([127.3, 121.6, 130.2]), axis=2)
    assert np.abs(rad - 60.0).max() < 0.05


def test_mesh_iso_above_zero_classifies_every_brick(o32):
    """iso > 0: the occupancy flags (which only know about values below zero)
    are not used; the result still matches."""
    print("case: mesh iso 0.05", flush=True)
    n = 64
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    r = np.sqrt((x - 31.7) ** 2 + (y - 32.2) ** 2 + (z - 30.9) ** 2)
    vol = np.zeros((n, n, n, 2), np.float32)
    vol[..., 0] = (r - 10.0) * 0.01
    vol[..., 1] = 1
    g = GFusion()
    g.setConfig(kinect_config(n, 2.0, "f32"))
    g.allocateBuffers()
    g.setVolume(*planes(vol, "f32"))
    _, V = gpu_mesh(g, iso=0.05)
    W = o32.mc_extract(vol, oracle.F32X2, iso=0.05)
    assert len(W) > 100
    np.testing.assert_array_equal(V, W)


def test_stl_and_ply_files(o32, seq20, tmp_path):
    print("case: STL / PLY export", flush=True)
    depth, gt = seq20
    g = GFusion()
    g.setConfig(kinect_config(128, 2.0, "rg16i"))
    g.allocateBuffers()
    g.Reset(gt[0])
    for f in range(4):
        g.frame(depth[f].astype(np.float32))
    mc, V = gpu_mesh(g)
    nt = len(V)
    assert nt > 0
    stl = tmp_path / "mesh.stl"
    mc.exportMesh(str(stl))
    data = stl.read_bytes()
    assert len(data) == 84 + 50 * nt
    assert data[:80] == bytes(80)
    assert struct.unpack("<I", data[80:84])[0] == 4 * nt  # posData.size() / 3 (mcubes.cpp:445)
    rec = np.frombuffer(data[84:], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    assert np.all(rec["n"] == 1.0) and np.all(rec["a"] == 0)
    np.testing.assert_array_equal(rec["v"], V[:, :, :3])
    ply = tmp_path / "mesh.ply"
    mc.exportPointCloud(str(ply))
    lines = ply.read_text().splitlines()
    assert lines[0] == "ply" and lines[2] == "element vertex %d" % (3 * nt) and lines[6] == "end_header"
    pts = np.array([[float(t) for t in l.split()] for l in lines[7:]])
    np.testing.assert_allclose(pts, V[:, :, :3].reshape(-1, 3), rtol=1e-5, atol=1e-4)


def test_mesh_512_after_frames_is_consistent(seq20):
    """512^3 (the benchmark volume) after a few frames: every vertex lies on a
    cube edge inside the 10-voxel border, the triangle count is stable across
    repeated calls and the vertex buffer is identical."""
    print("case: mesh 512^3 f32", flush=True)
    depth, gt = seq20
    g = GFusion()
    g.setConfig(kinect_config(512, 2.0, "f32"))
    g.allocateBuffers()
    g.Reset(gt[0])
    for f in range(6):
        g.frame(depth[f].astype(np.float32))
    _, V = gpu_mesh(g)
    _, V2 = gpu_mesh(g)
    assert len(V) > 10000
    np.testing.assert_array_equal(V, V2)
    P = V[:, :, :3]
    # mix() in float can land an ulp below the cube's first coordinate
    assert P.min() >= 11 - 1e-4 and P.max() <= 512 - 9 + 1e-4
    near = np.abs(P - np.round(P)) < 1e-4
    assert np.all(near.sum(axis=2) >= 2)  # on an edge: two integer coordinates


@pytest.mark.parametrize("storage", ["rg16i", "f32", "f16"])
def test_intensity_projection_matches_oracle(o32, seq20, storage):
    """intensityProjection (mip.cs) on a fused 128^3 volume: the same FP32
    sample expressions and additions on both sides.  The GPU splits each ray
    over four lanes and adds their samples back in march order, and skips
    samples in bricks that hold only (0, 0): neither may change a bit, with
    the skipping on or off."""
    print("case: intensity projection 128^3 %s" % storage, flush=True)
    depth, gt = seq20
    cfg = kinect_config(128, 2.0, storage)
    s = oracle.OracleSession(o32, 128, 2.0, cfg.mu, STORAGE_ORACLE[storage])
    s.reset(gt[0])
    for f in range(5):
        s.frame(depth[f].astype(np.float32))
    vol = s.volume()
    g = GFusion()
    g.setConfig(cfg)
    g.allocateBuffers()
    g.Reset(gt[0])
    g.setVolume(*planes(vol, storage))
    g.setPose(s.pose())
    _, zero = g.getBrickFlags()
    assert 0.1 < (zero == 0).mean() < 0.9  # bricks of (0, 0) to skip, and others
    img = g.intensityProjection()
    want = o32.intensity_projection(vol, STORAGE_ORACLE[storage], (2.0,) * 3, s.pose(), (365.5, 365.5, 255.5, 211.5),
                                    0.1, 3.0, np.float32(0.75) * np.float32(cfg.mu))
    assert (want[..., 0] > 0).mean() > 0.2
    np.testing.assert_array_equal(img, want)
    g.setOption(K.OPT_SKIP_EMPTY, 0)
    np.testing.assert_array_equal(g.intensityProjection(), want)
