"""Host-side checks that need no GPU: the C ABI library loads and exports every
entry point include/kfusion.h declares, configuration defaults and
validation, admission sizes, and the synthetic sequence generator."""
import ctypes as C

import numpy as np
import pytest

from oglfusionkinectv2_amd import _lib as K
from oglfusionkinectv2_amd import GFusionConfig, kinect_config, synth
from oglfusionkinectv2_amd.gfusion import to_kf_config


def test_library_exports_every_declared_symbol():
    L = K.lib()
    names = K.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.kf_version().startswith(b"kfusion")


def test_config_defaults_match_gfusionconfig():
    c = K.KfConfig()
    assert K.lib().kf_config_default(C.byref(c)) == 0
    ref = GFusionConfig()  # include/gFusion.h:55-70
    assert tuple(c.volume_size) == ref.volumeSize
    assert tuple(c.volume_dim) == pytest.approx(ref.volumeDimensions)
    assert c.mu == pytest.approx(ref.mu) and c.max_weight == ref.maxWeight
    assert (c.near_plane, c.far_plane) == pytest.approx((ref.nearPlane, ref.farPlane))
    assert (c.dist_threshold, c.normal_threshold) == pytest.approx((ref.dist_threshold, ref.normal_threshold))
    assert list(c.iterations) == ref.iterations
    assert (c.width, c.height) == (512, 424)


def test_step_size_and_mu_of_benchmark_configs():
    c = kinect_config(512, 2.0, "f32")
    assert c.stepSize() == np.float32(2.0) / np.float32(512)
    assert c.mu == pytest.approx(3 * 2.0 / 512, rel=1e-6)
    assert kinect_config(128, 2.0).mu == pytest.approx(0.046875)
    assert c.iterations == [4, 5, 10]  # 10/5/4 coarse -> fine


@pytest.mark.parametrize("storage,bv", [("rg16i", 4), ("f32", 8), ("f16", 4)])
def test_session_bytes(storage, bv):
    kc = to_kf_config(kinect_config(512, 2.0, storage))
    n = C.c_size_t()
    assert K.lib().kf_session_bytes(C.byref(kc), C.byref(n)) == 0
    vol = 512 ** 3 * bv
    assert vol <= n.value <= vol + 40 * 1024 * 1024


@pytest.mark.parametrize("field,value", [
    ("volume_size", (0, 64, 64)), ("volume_size", (4096, 64, 64)), ("mu", 0.0), ("mu", -1.0),
    ("max_weight", 0.0), ("near_plane", 5.0), ("width", 510), ("height", 0), ("storage", 7),
    ("reduce_rule", 3), ("iterations", (-1, 5, 5)),
])
def test_invalid_configs_are_rejected(field, value):
    kc = to_kf_config(kinect_config(64, 2.0))
    if isinstance(value, tuple):
        for i, v in enumerate(value):
            getattr(kc, field)[i] = v
    else:
        setattr(kc, field, value)
    n = C.c_size_t()
    rc = K.lib().kf_session_bytes(C.byref(kc), C.byref(n))
    assert rc == K.KF_ERR_INVALID
    assert K.lib().kf_last_error()
    h = C.c_void_p()
    assert K.lib().kf_session_create(C.byref(kc), None, C.byref(h)) == K.KF_ERR_INVALID
    assert not h.value


def test_null_arguments_are_errors_not_crashes():
    L = K.lib()
    assert L.kf_session_destroy(None) == 0
    assert L.kf_raycast(None) == K.KF_ERR_INVALID
    assert L.kf_track(None, None) == K.KF_ERR_INVALID
    assert L.kf_integrate(None, 0) == K.KF_ERR_INVALID
    assert L.kf_config_default(None) == K.KF_ERR_INVALID


# ---------------------------------------------------------------- synthetic
def test_trajectory_bounds():
    P = synth.trajectory(1000, 300)
    np.testing.assert_array_equal(P[0], np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 0], [0, 0, 0, 1]], np.float32))
    dt = np.linalg.norm(np.diff(P[:, :3, 3], axis=0), axis=1)
    assert dt.max() <= 0.01 and dt.mean() > 0.002
    R = P[:, :3, :3].astype(np.float64)
    rel = np.einsum("nji,njk->nik", R[:-1], R[1:])
    ang = np.degrees(np.arccos(np.clip((np.trace(rel, axis1=1, axis2=2) - 1) / 2, -1, 1)))
    assert ang.max() <= 1.0
    np.testing.assert_allclose(np.einsum("nji,njk->nik", R, R), np.broadcast_to(np.eye(3), R.shape), atol=1e-6)


def test_depth_noise_dropout_and_range():
    d, P = synth.sequence(3, 2)
    clean = synth.render(1003, P, noise=False)
    assert d.dtype == np.uint16 and d.shape == (2, 424, 512)
    valid = clean > 0
    drop = valid & (d == 0)
    assert 0.01 < drop.sum() / valid.sum() < 0.03  # 2 % dropout
    both = valid & (d > 0)
    z = clean[both] / 1000.0
    r = (d[both].astype(np.float64) - clean[both]) / 1000.0
    sigma = 0.0012 + 0.0019 * (z - 0.4) ** 2
    assert abs(np.mean(r / sigma)) < 0.05 and 0.9 < np.std(r / sigma) < 1.15
    assert d.max() <= 4500
    np.testing.assert_array_equal(d, synth.sequence(3, 2)[0])  # deterministic


def test_clean_depth_is_exact_intersection():
    P = synth.trajectory(1000, 1)
    d = synth.render(1000, P, noise=False)[0].astype(np.float64) / 1000
    y, x = np.nonzero(d > 0)
    pts = np.stack([(x - 255.5) / 365.5 * d[y, x], (y - 211.5) / 365.5 * d[y, x], d[y, x]], -1)
    pts = pts + P[0][:3, 3]
    sdf = synth.scene_sdf(0, pts)
    assert np.percentile(np.abs(sdf), 99) < 1e-3  # within mm rounding of the surface
