"""The reference's own compute shaders on Mesa's llvmpipe (test infrastructure
only): a ctypes wrapper of oracle/_ref/libkfref.so, built by
oracle/build_ref.py from the reference tree.  It checks the C oracle against
the shaders themselves, GLSL compiler and texture plumbing included
(glGenerateMipmap of the R32F depth texture, RG16I image load/store
conversions).

Arrays use the oracle's layouts: maps (H, W, 4) float32, depth (H, W)
float32 in mm, the TSDF (Z, Y, X, 2) int16, poses 4x4 row-major (passed to
GL column-major, as glm does), track records as planes (result int32, error
float32, J (N, 6) float32).

All calls run on the one GL context of the process (llvmpipe renders with its
own thread pool; LP_NUM_THREADS sets its size).
"""
import ctypes as C
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libkfref.so")

_lib = None


def built():
    """True when libkfref.so exists (nothing is loaded)."""
    return os.path.exists(LIB)


def available():
    """True when libkfref.so is built and its context and shaders come up
    (loads Mesa into this process)."""
    try:
        lib()
        return True
    except (OSError, RuntimeError):
        return False


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise OSError("%s is not built (python oracle/build_ref.py)" % LIB)
        L = C.CDLL(LIB)
        L.kfr_last_error.restype = C.c_char_p
        L.kfr_info.restype = C.c_char_p
        if L.kfr_init() != 0:
            raise RuntimeError("llvmpipe reference context: %s" % L.kfr_last_error().decode())
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _cm(M):
    """4x4 row-major -> column-major float32 (glm's memory order)."""
    return _f(np.asarray(M, dtype=np.float64).T)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().kfr_last_error().decode())


class RefShaders:
    """gFusion's textures and buffers for one depth size and volume size."""

    def __init__(self, w=512, h=424, n=(64, 64, 64)):
        self.w, self.h = w, h
        self.n = tuple(int(v) for v in n)
        self.L = lib()
        _check(self.L.kfr_setup(w, h, *self.n))

    @staticmethod
    def info():
        return lib().kfr_info().decode()

    # -- pyramid ----------------------------------------------------------------
    def depth_to_vertex(self, depth_mm, cam):
        """Uploads level 0, glGenerateMipmap, depthToVert.cs on 3 levels."""
        d = _f(depth_mm)
        assert d.shape == (self.h, self.w)
        _check(self.L.kfr_depth_to_vertex(_p(d), _p(_f(cam))))

    def vertex_to_normal(self):
        _check(self.L.kfr_vertex_to_normal())

    def map(self, which, level=0):
        """which: 'depth', 'vertex', 'normal', 'ref_vertex', 'ref_normal'."""
        k = ["depth", "vertex", "normal", "ref_vertex", "ref_normal"].index(which)
        h, w = self.h >> level, self.w >> level
        out = np.zeros((h, w) if k == 0 else (h, w, 4), np.float32)
        _check(self.L.kfr_read_map(k, level, _p(out)))
        return out

    def set_map(self, which, arr, level=0):
        k = ["depth", "vertex", "normal", "ref_vertex", "ref_normal"].index(which)
        a = _f(arr)
        assert a.shape[:2] == (self.h >> level, self.w >> level)
        _check(self.L.kfr_write_map(k, level, _p(a)))

    # -- ICP ---------------------------------------------------------------------
    def track(self, layer, view, pose, dist_th=0.05, normal_th=0.8):
        """track.cs at `layer` (gFusion::track, gFusion.cpp:742-777): `view`
        is K * inverse(m_pose), Ttrack is m_pose (row-major 4x4 each)."""
        _check(self.L.kfr_track(layer, _p(_cm(view)), _p(_cm(pose)), C.c_float(dist_th), C.c_float(normal_th)))

    def track_records(self, n):
        res = np.zeros(n, np.int32)
        err = np.zeros(n, np.float32)
        J = np.zeros((n, 6), np.float32)
        _check(self.L.kfr_read_track(n, _p(res), _p(err), _p(J)))
        return res, err, J

    def set_track_records(self, res, err, J):
        res = np.ascontiguousarray(res, np.int32)
        err, J = _f(err), _f(J)
        _check(self.L.kfr_write_track(len(res), _p(res), _p(err), _p(J)))

    def reduce(self, layer):
        """reduce.cs over the records of level `layer`: the 8 x 32 partials."""
        out = np.zeros(256, np.float32)
        _check(self.L.kfr_reduce(layer, _p(out)))
        return out.reshape(8, 32)

    # -- point-to-volume tracking (TrackSDF) ---------------------------------------
    def track_sdf(self, layer, cam_to_world, dim):
        """trackSDF.cs at `layer` (gFusion::trackSDF, gFusion.cpp:1040-1098)
        with Ttrack = camToWorld; the records stay in the SSBO."""
        _check(self.L.kfr_track_sdf(layer, _p(_cm(cam_to_world)), _p(_f(dim))))

    def sdf_records(self):
        """The (h*w, 8) record buffer {h, D, J[6]} of all level-0 pixels."""
        out = np.zeros((self.h * self.w, 8), np.float32)
        _check(self.L.kfr_read_sdf(self.h * self.w, _p(out)))
        return out

    def set_sdf_records(self, rec):
        r = _f(rec).reshape(-1, 8)
        _check(self.L.kfr_write_sdf(len(r), _p(r)))

    def reduce_sdf(self, layer):
        """reduceSDF.cs over level `layer`: the 8 x 32 partials."""
        out = np.zeros(256, np.float32)
        _check(self.L.kfr_reduce_sdf(layer, _p(out)))
        return out.reshape(8, 32)

    def intensity_projection(self, view, near, far, step, large_step, dim):
        """mip.cs (gFusion::intensityProjection, gFusion.cpp:1368-1409): the
        (h, w, 4) image; `view` = m_pose * invK."""
        out = np.zeros((self.h, self.w, 4), np.float32)
        _check(self.L.kfr_intensity_projection(_p(_cm(view)), C.c_float(near), C.c_float(far), C.c_float(step),
                                               C.c_float(large_step), _p(_f(dim)), _p(out)))
        return out

    # -- volume ------------------------------------------------------------------
    def set_volume(self, vol):
        v = np.ascontiguousarray(vol, np.int16)
        assert v.shape == (self.n[2], self.n[1], self.n[0], 2)
        _check(self.L.kfr_write_volume(_p(v)))

    def volume(self):
        v = np.zeros((self.n[2], self.n[1], self.n[0], 2), np.int16)
        _check(self.L.kfr_read_volume(_p(v)))
        return v

    def integrate(self, inv_track, K, mu, dim, max_weight=100.0):
        """integrate.cs (gFusion::integrate, gFusion.cpp:1234-1287):
        `inv_track` = inverse(m_pose), `K` = getCameraMatrix(camPams); the
        depth texture's level 0 (the last uploaded frame) is its input."""
        _check(self.L.kfr_integrate(_p(_cm(inv_track)), _p(_cm(K)), C.c_float(mu), C.c_float(max_weight),
                                    _p(_f(dim))))

    def raycast(self, view, near, far, step, large_step, dim):
        """raycast.cs (gFusion::raycast, gFusion.cpp:1289-1366): `view` =
        m_pose * invK; fills the reference vertex / normal maps."""
        _check(self.L.kfr_raycast(_p(_cm(view)), C.c_float(near), C.c_float(far), C.c_float(step),
                                  C.c_float(large_step), _p(_f(dim))))


class RefSession:
    """gFusion's frame loop with the reference's shaders doing the per-pixel
    and per-voxel work and the oracle's restatement of the host code (glm
    matrices, getReduction, the 6x6 solve, the Euler pose update) around them:
    depthToVertex + vertexToNormal, Track() (gFusion.cpp:648-740), integrate
    while tracked or within the first frames after Reset (kinect.cpp:654-695),
    raycast; TrackSDF (gFusion.cpp:835-1038), the pose library and
    recoverPose (:1570-1653) as the failure path.

    `times` accumulates wall-clock seconds per phase (each phase ends with
    glFinish)."""

    PHASES = ("pyramid", "track", "integrate", "raycast")

    def __init__(self, volume, dim, mu, iterations=(4, 5, 10), cam=(365.5, 365.5, 255.5, 211.5), width=512,
                 height=424, rule=0, max_weight=100.0, near=0.1, far=3.0, dist_th=0.05, normal_th=0.8):
        from . import Oracle
        self.o = Oracle("f32")
        n = volume if isinstance(volume, (tuple, list)) else (volume,) * 3
        self.N = tuple(int(v) for v in n)
        self.g = RefShaders(width, height, self.N)
        self.dim = (float(dim),) * 3 if np.isscalar(dim) else tuple(float(v) for v in dim)
        self.mu, self.iterations, self.cam = float(mu), tuple(iterations), tuple(float(c) for c in cam)
        self.w, self.h, self.rule, self.max_weight = width, height, rule, float(max_weight)
        self.near, self.far, self.dist_th, self.normal_th = float(near), float(far), float(dist_th), float(normal_th)
        # stepSize() (gFusion.h:72-81) and largeStep = 0.75 * mu (gFusion.cpp:1304)
        d32 = np.float32(min(self.dim))
        self.step = float(d32 / np.float32(max(self.N)))
        self.large_step = float(np.float32(0.75) * np.float32(self.mu))
        self.K = self.o.camera_matrix(self.cam)
        self.invK = self.o.inverse_camera_matrix(self.cam)
        self.pose_ = np.eye(4, dtype=np.float32)
        self.pose_eig = np.eye(4, dtype=np.float32)
        self.cum_twist = np.zeros(6)
        self.library = []
        self.track_mode, self.recover = 1, 1  # kinect.cpp's defaults
        # glBufferData(..., NULL) (gFusion.cpp:349): contents taken as zero
        self.g.set_sdf_records(np.zeros((width * height, 8), np.float32))
        self.counter, self.reset_flag = 0, 1
        self.ae, self.tracked, self.iters_run = 0.0, 0, 0
        self.times = dict.fromkeys(self.PHASES, 0.0)

    def _t(self, phase, t0):
        self.g.L.kfr_finish()
        self.times[phase] += time.perf_counter() - t0

    def reset(self, pose):
        """Reset(pose, false) + raycast() (gFusion.cpp:226-287, kinect.cpp:472-474):
        resetVolume zeroes every texel."""
        self.g.set_volume(np.zeros((self.N[2], self.N[1], self.N[0], 2), np.int16))
        self.pose_ = np.asarray(pose, np.float32).copy()
        # resetPose (gFusion.cpp:275-286): m_pose_eig = identity rotation with
        # the pose's translation, m_cumTwist = 0, the library emptied
        self.pose_eig = np.eye(4, dtype=np.float32)
        self.pose_eig[:3, 3] = self.pose_[:3, 3]
        self.cum_twist = np.zeros(6)
        self.library = []
        self.raycast()
        self.counter, self.reset_flag = 0, 1

    def pose(self):
        return self.pose_.copy()

    def set_pose(self, pose):
        self.pose_ = np.asarray(pose, np.float32).copy()

    def depth(self, depth_mm):
        t0 = time.perf_counter()
        self.g.depth_to_vertex(depth_mm, self.cam)
        self.g.vertex_to_normal()
        self._t("pyramid", t0)

    def track(self):
        """Track(): levels 2 -> 0, iterations[level] each, early exit on
        |x| < 1e-5 (x != 0); accepted iff 0 < AE <= 0.5 of the last iteration."""
        t0 = time.perf_counter()
        old = self.pose_.copy()
        ae = np.float32(0)
        self.iters_run = 0
        for level in (2, 1, 0):
            for _ in range(self.iterations[level]):
                view = self.o.track_view(self.pose_, self.cam)
                self.g.track(level, view, self.pose_, self.dist_th, self.normal_th)
                partials = self.g.reduce(level)
                b, Cm, ae = self.o.get_reduction(partials, self.rule)
                x = self.o.solve6(Cm, b)
                self.pose_ = self.o.pose_update(x, self.pose_).astype(np.float32)
                self.iters_run += 1
                nrm = float(np.sqrt(np.dot(x, x)))
                if nrm < 1e-5 and nrm != 0:
                    break
        if ae > np.float32(0.5) or ae == 0:
            self.pose_, self.tracked = old, 0
        else:
            self.tracked = 1
            self._update_pose_finder()
        self.ae = float(ae)
        self._t("track", t0)
        return self.tracked

    def _update_pose_finder(self):
        """updatePoseFinder (gFusion.cpp:1570-1597): keep the pose when its
        translation is at least 0.2 m from every kept pose's."""
        t = self.pose_[:3, 3]
        for q in self.library:
            d = (t - q[:3, 3]).astype(np.float32)
            if np.sqrt(np.float32(np.dot(d, d))) < np.float32(0.2):
                return
        if len(self.library) < 256:
            self.library.append(self.pose_.copy())

    def track_sdf(self):
        """TrackSDF() from m_pose_eig: per level 2 -> 0 and iteration,
        camToWorld = exp(twist(result)) * m_pose_eig_prev, trackSDF +
        reduceSDF + getSDFReduction, a scaled and damped LDLT step; a converged
        iteration ends its level and marks the frame tracked."""
        t0 = time.perf_counter()
        old, prev = self.pose_.copy(), self.pose_eig.copy()
        result, result_prev = np.zeros(6), np.zeros(6)
        tracked = 0
        for level in (2, 1, 0):
            for it in range(self.iterations[level]):
                cam_to_world = self.o.mul4(self.o.twist_exp(result), prev)
                self.g.track_sdf(level, cam_to_world, self.dim)
                partials = self.g.reduce_sdf(level)
                b, Cm, ae = self.o.get_reduction(partials, self.rule)
                cn, result, result_prev = self.o.sdf_step(b, Cm, it, result, result_prev)
                if (float(ae) != 0.0 and float(ae) < 1e-3) or cn < 1e-3:
                    tracked = 1
                    break
        if np.isnan(result.sum()):
            result = np.zeros(6)
        if tracked:
            self.pose_eig = self.o.mul4(self.o.twist_exp(result), prev).astype(np.float32)
            self.pose_ = self.pose_eig.copy()
            self.pose_[3, :3] = 0
            self.pose_[3, 3] = 1
            self.cum_twist = self.cum_twist + result
        else:
            self.pose_eig, self.pose_ = prev, old
        self.tracked = tracked
        self._t("track", t0)
        return tracked

    def recover_pose(self):
        """recoverPose() (gFusion.cpp:1616-1653): each kept pose in turn as
        m_pose, then TrackSDF(); stop at the first success."""
        for q in self.library:
            self.pose_ = q.copy()
            if self.track_sdf():
                return 1
        return 0

    def set_mode(self, track_mode=1, recover=1):
        """track_mode: 1 = Track, 2 = TrackSDF, 3 = both; recover: call
        recoverPose() on a frame whose tracking failed."""
        self.track_mode, self.recover = int(track_mode), int(recover)

    def integrate(self):
        t0 = time.perf_counter()
        self.g.integrate(self.o.inverse4(self.pose_), self.K, self.mu, self.dim, self.max_weight)
        self._t("integrate", t0)

    def raycast(self):
        t0 = time.perf_counter()
        view = self.o.mul4(self.pose_, self.invK)
        self.g.raycast(view, self.near, self.far, self.step, self.large_step, self.dim)
        self._t("raycast", t0)

    def frame(self, depth_mm):
        """The frame block of kinect.cpp:654-695 with integration on."""
        self.depth(depth_mm)
        tracked = 0
        if self.track_mode & 1:
            tracked = self.track()
        if self.track_mode & 2:
            tracked = self.track_sdf()
        self.tracked = tracked
        if tracked or self.reset_flag:
            self.integrate()
            if self.counter > 2:
                self.reset_flag = 0
        if not tracked and self.recover:
            self.recover_pose()
        self.counter += 1
        self.raycast()
        return tracked

    def volume(self):
        return self.g.volume()

    def map(self, which, level=0):
        return self.g.map(which, level)
