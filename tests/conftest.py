import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def o32():
    import oracle
    return oracle.Oracle("f32")


@pytest.fixture(scope="session")
def o64():
    import oracle
    return oracle.Oracle("f64")


@pytest.fixture(scope="session")
def seq20():
    """20 frames of synthetic session 0: (depth u16 [20,424,512], gt poses)."""
    from oglfusionkinectv2_amd import synth
    return synth.sequence(0, 20)


def small_cam(w, h):
    """Kinect v2 intrinsics scaled to a (w, h) image."""
    f = 365.5 * w / 512.0
    return (f, f, w / 2.0 - 0.5, h / 2.0 - 0.5)


def rot(ax, ay, az):
    from scipy.spatial.transform import Rotation
    return Rotation.from_euler("XYZ", [ax, ay, az]).as_matrix()


def rigid(ax, ay, az, t):
    M = np.eye(4)
    M[:3, :3] = rot(ax, ay, az)
    M[:3, 3] = t
    return M.astype(np.float32)
