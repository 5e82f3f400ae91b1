"""Pins the oracle's intensityProjection (gFusion.cpp:1368-1409, mip.cs):
a constant positive volume sums to value x samples along each ray (the ray's
length inside the box / near-far range over the step), zero and negative
volumes give zero, and the image holds (sum, 0, 0, 1).  No GPU."""
import numpy as np

import oracle

CAM = (365.5, 365.5, 255.5, 211.5)


def _pose():
    P = np.eye(4, dtype=np.float32)
    P[:3, 3] = (1.0, 1.0, -0.5)
    return P


def test_constant_volume_sums_samples(o64):
    n, dim, ls = 32, 2.0, 0.05
    vol = np.zeros((n, n, n, 2))
    vol[..., 0] = 0.1
    img = o64.intensity_projection(vol, oracle.F64X2, (dim,) * 3, _pose(), CAM, 0.1, 3.0, ls, w=64, h=48)
    assert np.all(img[..., 1] == 0) and np.all(img[..., 2] == 0) and np.all(img[..., 3] == 1)
    # the centre ray (pixel (255, 211), direction ~ +z) from z = -0.5: box
    # entry at t = 0.5, exit at t = 2.5, so about (2.5 - 0.5) / ls samples
    img_full = o64.intensity_projection(vol, oracle.F64X2, (dim,) * 3, _pose(), CAM, 0.1, 3.0, ls, w=512, h=424)
    k = (2.5 - 0.5) / ls
    assert abs(img_full[211, 255, 0] - 0.1 * k) <= 0.1 * 1.01


def test_zero_and_negative_volumes(o32):
    n = 32
    zero = np.zeros((n, n, n, 2), np.float32)
    neg = zero.copy()
    neg[..., 0] = -0.05
    for vol in (zero, neg):
        img = o32.intensity_projection(vol, oracle.F32X2, (2.0,) * 3, _pose(), CAM, 0.1, 3.0, 0.05, w=64, h=48)
        assert np.all(img[..., 0] == 0) and np.all(img[..., 3] == 1)


def test_rg16i_scales_the_samples(o32):
    n = 32
    q = np.zeros((n, n, n, 2), np.int16)
    q[..., 0] = 16383
    f = np.zeros((n, n, n, 2), np.float32)
    f[..., 0] = np.float32(16383) * np.float32(0.00003051944088)
    a = o32.intensity_projection(q, oracle.RG16I, (2.0,) * 3, _pose(), CAM, 0.1, 3.0, 0.05, w=64, h=48)
    b = o32.intensity_projection(f, oracle.F32X2, (2.0,) * 3, _pose(), CAM, 0.1, 3.0, 0.05, w=64, h=48)
    np.testing.assert_allclose(a, b, rtol=1e-5)
