"""Seeded synthetic depth sequences (ctypes over libkfsynth.so).

Scene = inside-out 4x3x4 m room + 12 spheres + 6 oriented boxes (seed =
1000 + session id by convention); trajectory = smooth periodic spline, at most
1 cm / 1 degree per frame, starting at translate(1, 1, 0); depth = 512x424
uint16 mm with the Kinect v2 noise model (see csrc/kf_synth.cpp).
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libkfsynth.so")
CAM = (365.5, 365.5, 255.5, 211.5)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libkfsynth.so is not built; run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        L.kfs_trajectory.argtypes = [C.c_uint64, C.c_int, C.c_void_p]
        L.kfs_render.argtypes = [C.c_uint64, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                 C.c_int, C.c_void_p]
        L.kfs_scene_sdf.argtypes = [C.c_uint64, C.c_void_p, C.c_long, C.c_void_p]
        for f in (L.kfs_trajectory, L.kfs_render, L.kfs_scene_sdf):
            f.restype = C.c_int
        _lib = L
    return _lib


def trajectory(seed, frames):
    """(frames, 4, 4) float32 camera->world poses, row-major."""
    cm = np.zeros((frames, 16), dtype=np.float32)
    if lib().kfs_trajectory(seed, frames, cm.ctypes.data) != 0:
        raise RuntimeError("kfs_trajectory failed")
    return np.ascontiguousarray(cm.reshape(frames, 4, 4).transpose(0, 2, 1))


def render(seed, poses, first_frame=0, noise=True, cam=CAM, width=512, height=424):
    """uint16 mm depth of shape (frames, height, width) for row-major poses."""
    poses = np.asarray(poses, dtype=np.float32).reshape(-1, 4, 4)
    cm = np.ascontiguousarray(poses.transpose(0, 2, 1)).reshape(-1, 16)
    out = np.zeros((len(poses), height, width), dtype=np.uint16)
    camv = np.asarray(cam, dtype=np.float32)
    if lib().kfs_render(seed, cm.ctypes.data, len(poses), first_frame, camv.ctypes.data, width, height,
                        1 if noise else 0, out.ctypes.data) != 0:
        raise RuntimeError("kfs_render failed")
    return out


def sequence(session_id, frames, noise=True):
    """(depth u16 [F,424,512], poses [F,4,4]) of session `session_id`."""
    seed = 1000 + session_id
    poses = trajectory(seed, frames)
    return render(seed, poses, noise=noise), poses


def scene_sdf(session_id, points):
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(len(p), dtype=np.float32)
    if lib().kfs_scene_sdf(1000 + session_id, p.ctypes.data, len(p), out.ctypes.data) != 0:
        raise RuntimeError("kfs_scene_sdf failed")
    return out
