"""Input adapters around the frame loop (host side).

loadCalibration mirrors kinect.cpp:78-161 through the library's parser
(kf_load_calibration): OpenCV-style XML or JSON depth-camera calibrations.
Recorded depth is read as raw frames of the sensor's layout (512 x 424,
float mm as libfreenect2's undistorted frame, interface.cpp:311-329, or
uint16 mm) to feed GFusion.depthToVertex / the _dev upload paths.
"""
import ctypes as C

import numpy as np

from . import _lib as K


class kf_cam_params(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("fx", "fy", "ppx", "ppy", "k1", "k2", "p1", "p2", "k3")]


class CamPams:
    """camPams (kinect.h): intrinsics fx, fy, ppx, ppy and distortion k1, k2, p1, p2, k3."""

    def __init__(self, **kw):
        for n, _ in kf_cam_params._fields_:
            setattr(self, n, float(kw.get(n, 0.0)))

    def depth_cam(self):
        """The vec4 the frame loop hands to setCameraParams (kinect.cpp:176-177):
        (fx, fx, ppx, ppy) -- fx twice, as the reference does."""
        return (self.fx, self.fx, self.ppx, self.ppy)


def loadCalibration(path):
    """Returns (1, CamPams) on success; (0, None) for an extension other than
    .xml / .json (the reference prints and returns 0); raises KfError for a
    file it cannot read or parse."""
    p = str(path)
    if not (p.endswith(".xml") or p.endswith(".json")):
        print("please use a .xml or .json file for loading calibration data")
        return 0, None
    c = kf_cam_params()
    K.check(K.lib().kf_load_calibration(p.encode(), C.byref(c)))
    return 1, CamPams(**{n: getattr(c, n) for n, _ in kf_cam_params._fields_})


def read_depth_frames(path, width=512, height=424, dtype=np.float32):
    """Recorded depth: consecutive raw frames of height x width `dtype`
    (float32 or uint16 millimetres), memory-mapped as (frames, height, width)."""
    dt = np.dtype(dtype)
    raw = np.memmap(path, dtype=dt, mode="r")
    n = raw.size // (width * height)
    return raw[:n * width * height].reshape(n, height, width)


def write_depth_frames(path, frames):
    np.ascontiguousarray(frames).tofile(path)
