"""Host-side mirror of the reference's `class gFusion` (include/gFusion.h:124-276).

Method names, argument meaning and call order follow the reference so that a
frame loop written against gFusion (kinect.cpp:654-695) reads the same here:

    g = GFusion()
    g.setConfig(cfg); g.setCameraParams((fx, fy, cx, cy)); g.setPose(init)
    g.initTextures(); g.initVolume(); g.allocateBuffers()
    g.Reset(init, False); g.raycast()
    for depth in frames:
        g.depthToVertex(depth); g.vertexToNormal()
        tracked = g.Track()
        if tracked or reset_window: g.integrate()
        g.raycast()

GL texture ids are replaced by device buffers of a kfusion session; the
getters that returned texture ids return host copies instead.  All compute runs
in the HIP library (libkfusion.so) -- there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _lib as K


class GFusionConfig:
    """gFusionConfig (include/gFusion.h:42-82) with the reference defaults."""

    def __init__(self):
        self.volumeSize = (256, 256, 256)
        self.volumeDimensions = (1.0, 1.0, 1.0)
        self.mu = 0.05
        self.maxWeight = 100.0
        self.nearPlane = 0.1
        self.farPlane = 3.0
        self.dist_threshold = 0.05
        self.normal_threshold = 0.8
        self.iterations = [5, 5, 5]
        self.track_threshold = 0.5
        # not part of gFusionConfig: storage format and ICP combine rule
        self.storage = "rg16i"
        self.reduce_rule = K.RULE_LAST_ROW

    def stepSize(self):
        """include/gFusion.h:72-81"""
        return np.float32(min(self.volumeDimensions)) / np.float32(max(self.volumeSize))


def kinect_config(volume=512, dim=2.0, storage="f32", mu_voxels=3.0, iterations=(4, 5, 10),
                  reduce_rule=K.RULE_LAST_ROW):
    """The synthetic-sequence configuration: N^3 voxels over a `dim` cube,
    truncation `mu_voxels` voxels, ICP iterations by level (0 = finest)."""
    c = GFusionConfig()
    n = volume if isinstance(volume, (tuple, list)) else (volume,) * 3
    c.volumeSize = tuple(int(v) for v in n)
    c.volumeDimensions = (float(dim),) * 3
    c.mu = float(np.float32(dim) / np.float32(max(n)) * np.float32(mu_voxels))
    c.iterations = list(iterations)
    c.storage = storage
    c.reduce_rule = reduce_rule
    return c


KINECT_V2_CAM = (365.5, 365.5, 255.5, 211.5)


def to_kf_config(cfg, cam=KINECT_V2_CAM, width=512, height=424):
    kc = K.KfConfig()
    K.check(K.lib().kf_config_default(C.byref(kc)))
    for i in range(3):
        kc.volume_size[i] = int(cfg.volumeSize[i])
        kc.volume_dim[i] = float(cfg.volumeDimensions[i])
        kc.iterations[i] = int(cfg.iterations[i])
    kc.mu = cfg.mu
    kc.max_weight = cfg.maxWeight
    kc.near_plane = cfg.nearPlane
    kc.far_plane = cfg.farPlane
    kc.dist_threshold = cfg.dist_threshold
    kc.normal_threshold = cfg.normal_threshold
    for i in range(4):
        kc.cam[i] = float(cam[i])
    kc.width = width
    kc.height = height
    kc.storage = K.STORAGE_NAMES[cfg.storage] if isinstance(cfg.storage, str) else int(cfg.storage)
    kc.reduce_rule = int(cfg.reduce_rule)
    return kc


def translate(x, y, z):
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = (x, y, z)
    return m


def _pose_arg(pose):
    """4x4 row-major numpy pose -> 16 column-major floats (glm layout)."""
    p = np.ascontiguousarray(np.asarray(pose, dtype=np.float32).T).reshape(16)
    return p, p.ctypes.data_as(C.POINTER(C.c_float))


class GFusion:
    """gFusion over a kfusion session.  Poses are 4x4 numpy arrays (row-major,
    camera -> world), converted to glm's column-major layout at the ABI."""

    def __init__(self, stream=None, size=(512, 424)):
        self.configuration = GFusionConfig()
        self._cam = KINECT_V2_CAM
        self._size = tuple(size)
        self._pose = np.eye(4, dtype=np.float32)
        self._stream = stream
        self._s = None

    # -- configuration (gFusion.h:180-195) ---------------------------------
    def setConfig(self, config):
        self.configuration = config

    def setCameraParams(self, camPams, camPamsColor=None):
        self._cam = tuple(float(v) for v in camPams)

    def setPose(self, pose):
        self._pose = np.asarray(pose, dtype=np.float32)
        if self._s:
            p, ptr = _pose_arg(self._pose)
            K.check(K.lib().kf_set_pose(self._s, ptr))

    def getPose(self):
        if not self._s:
            return self._pose.copy()
        out = np.zeros(16, dtype=np.float32)
        K.check(K.lib().kf_get_pose(self._s, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out.reshape(4, 4).T.copy()

    def getTransPose(self):
        """gFusion.h:267-274"""
        return self.getPose()[:3, 3].copy()

    # -- allocation (gFusion.cpp:200-330) ----------------------------------
    def initTextures(self):
        pass

    def initVolume(self):
        pass

    def allocateBuffers(self):
        """Creates the device session holding every buffer of the three init
        calls (textures, volume, reduction buffers)."""
        if self._s:
            K.check(K.lib().kf_session_destroy(self._s))
            self._s = None
        kc = to_kf_config(self.configuration, self._cam, *self._size)
        h = C.c_void_p()
        K.check(K.lib().kf_session_create(C.byref(kc), self._stream, C.byref(h)))
        self._s = h
        self.setPose(self._pose)

    @property
    def session(self):
        if not self._s:
            raise K.KfError(K.KF_ERR_STATE, "allocateBuffers() has not been called")
        return self._s

    def close(self):
        if self._s:
            K.lib().kf_session_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reset (gFusion.cpp:226-287) ---------------------------------------
    def Reset(self, pose, deleteFlag=False):
        """Zero the volume, set the pose; the caller's raycast() follows as in
        kinect.cpp:472-474 (kf_reset also raycasts and arms the reset window)."""
        if deleteFlag:
            self.allocateBuffers()
        self._pose = np.asarray(pose, dtype=np.float32)
        p, ptr = _pose_arg(self._pose)
        K.check(K.lib().kf_reset(self.session, ptr))

    def resetVolume(self):
        K.check(K.lib().kf_reset_volume(self.session))

    def resetPose(self, pose):
        self.setPose(pose)

    # -- per-frame calls ----------------------------------------------------
    def depthToVertex(self, depthArray):
        """gFusion.cpp:496: float mm (or uint16 mm) 512x424 depth; builds the
        depth mip chain, vertex and normal maps of all 3 levels."""
        d = np.ascontiguousarray(depthArray)
        if d.dtype == np.uint16:
            K.check(K.lib().kf_depth_to_vertex_u16(self.session, d.ctypes.data))
        else:
            d = np.ascontiguousarray(d, dtype=np.float32)
            K.check(K.lib().kf_depth_to_vertex_f32(self.session, d.ctypes.data))

    def vertexToNormal(self):
        """gFusion.cpp:528: the normal maps are produced by the same launch as
        the vertex maps in depthToVertex(); nothing is left to do here."""

    def Track(self):
        t = C.c_int(0)
        K.check(K.lib().kf_track(self.session, C.byref(t)))
        return bool(t.value)

    def TrackSDF(self):
        """gFusion.cpp:835-1038: point-to-volume tracking from m_pose_eig."""
        t = C.c_int(0)
        K.check(K.lib().kf_track_sdf(self.session, C.byref(t)))
        return bool(t.value)

    def recoverPose(self):
        """gFusion.cpp:1616-1653: TrackSDF from each pose-library entry."""
        t = C.c_int(0)
        K.check(K.lib().kf_recover_pose(self.session, C.byref(t)))
        return bool(t.value)

    def getPoseLibrary(self):
        """The poses of poseLibrary (gFusion.h:424) as (n, 4, 4) row-major."""
        n = C.c_int()
        buf = np.zeros((K.POSE_LIBRARY_MAX, 16), dtype=np.float32)
        K.check(K.lib().kf_read_pose_library(self.session, buf.ctypes.data_as(C.POINTER(C.c_float)),
                                             K.POSE_LIBRARY_MAX, C.byref(n)))
        return np.array([b.reshape(4, 4).T for b in buf[:n.value]]).reshape(n.value, 4, 4)

    def getCumTwist(self):
        """gFusion.h:260-265"""
        out = np.zeros(6, dtype=np.float32)
        K.check(K.lib().kf_get_cum_twist(self.session, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def setPoseEig(self, pose):
        """Loads m_pose_eig (a known tracking state; the reference sets it only
        in resetPose and TrackSDF)."""
        p, ptr = _pose_arg(pose)
        K.check(K.lib().kf_write_pose_eig(self.session, ptr))

    def sdfState(self):
        """m_pose_eig (row-major), cumulative twist, last TrackSDF partials /
        AE / iteration count and the per-iteration log rows."""
        pe = np.zeros(16, dtype=np.float32)
        tw = np.zeros(6, dtype=np.float32)
        parts = np.zeros(256, dtype=np.float32)
        ae, iters = C.c_float(), C.c_int()
        log = np.zeros((64, 10), dtype=np.float64)
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
        K.check(K.lib().kf_read_sdf_state(self.session, fp(pe), fp(tw), fp(parts), C.byref(ae), C.byref(iters),
                                          log.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(pose_eig=pe.reshape(4, 4).T.copy(), cum_twist=tw, partials=parts.reshape(8, 32),
                    ae=ae.value, iterations=iters.value, log=log[:min(iters.value, 64)].copy())

    def track(self, layer, first=True, save_pose=False):
        """One ICP step at `layer` (track + reduce + solve + update fused)."""
        K.check(K.lib().kf_icp_iteration(self.session, int(layer), int(first), int(save_pose)))

    def reduce(self, layer):
        """The reduction is fused into track(); kept for call compatibility."""

    def getReduction(self):
        """gFusion.cpp:1127-1168 on the last iteration's 8x32 partials:
        returns (b[6], C[6x6], alignmentEnergy) under the session's rule."""
        parts = self.readPartials()
        row = parts[7] if self.configuration.reduce_rule == K.RULE_LAST_ROW else parts.sum(0)
        b = row[1:7].copy()
        Cm = np.zeros((6, 6), dtype=np.float32)
        iu = np.triu_indices(6)
        Cm[iu] = row[7:28]
        Cm = Cm + np.triu(Cm, 1).T
        ae = np.float32(np.sqrt(np.float32(row[0]) / np.float32(row[28])))
        return b, Cm, ae

    def readPartials(self):
        out = np.zeros(256, dtype=np.float32)
        ae = C.c_float()
        K.check(K.lib().kf_read_reduction(self.session, out.ctypes.data_as(C.POINTER(C.c_float)),
                                          C.byref(ae)))
        return out.reshape(8, 32)

    def alignmentEnergy(self):
        ae = C.c_float()
        K.check(K.lib().kf_read_reduction(self.session, None, C.byref(ae)))
        return ae.value

    def intensityProjection(self):
        """gFusion.cpp:1368-1409: returns the (H, W, 4) image (sum of the
        positive TSDF samples per ray, 0, 0, 1)."""
        K.check(K.lib().kf_intensity_projection(self.session))
        return self._read_map(K.MAP_INTENSITY)

    def integrate(self, gated=False):
        K.check(K.lib().kf_integrate(self.session, K.INTEGRATE_GATED if gated else K.INTEGRATE_ALWAYS))

    def raycast(self):
        K.check(K.lib().kf_raycast(self.session))

    def frame(self, depthArray):
        """The frame block of kinect.cpp:654-695 without host synchronisation."""
        d = np.ascontiguousarray(depthArray, dtype=np.float32)
        K.check(K.lib().kf_frame_f32(self.session, d.ctypes.data))

    def status(self):
        t, it, c, r = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        K.check(K.lib().kf_read_status(self.session, C.byref(t), C.byref(it), C.byref(c), C.byref(r)))
        return dict(tracked=t.value, iterations=it.value, counter=c.value, reset=r.value)

    # -- read-back (gFusion.h:196-246) -------------------------------------
    def _read_map(self, which, level=0):
        w, h = self._size[0] >> level, self._size[1] >> level
        comps = 1 if which == K.MAP_DEPTH else 4
        out = np.zeros((h, w, comps), dtype=np.float32)
        K.check(K.lib().kf_read_map(self.session, which, level, out.ctypes.data, out.nbytes))
        return out if comps > 1 else out[..., 0]

    def _write_map(self, which, arr, level=0):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        K.check(K.lib().kf_write_map(self.session, which, level, a.ctypes.data, a.nbytes))

    def getVerts(self):
        """Reference (raycast) vertex map -- gFusion.h:210-215"""
        return self._read_map(K.MAP_REF_VERTEX)

    def getNorms(self):
        """Reference (raycast) normal map -- gFusion.h:216-222"""
        return self._read_map(K.MAP_REF_NORMAL)

    def getDepthImage(self):
        return self._read_map(K.MAP_DEPTH)

    def getVertexMap(self, level=0):
        return self._read_map(K.MAP_VERTEX, level)

    def getNormalMap(self, level=0):
        return self._read_map(K.MAP_NORMAL, level)

    def setReferenceMaps(self, verts, norms):
        self._write_map(K.MAP_REF_VERTEX, verts)
        self._write_map(K.MAP_REF_NORMAL, norms)

    def _plane_dtype(self):
        s = self.configuration.storage
        s = K.STORAGE_NAMES[s] if isinstance(s, str) else s
        return {K.STORAGE_RG16I: np.int16, K.STORAGE_F32: np.float32, K.STORAGE_F16: np.float16}[s]

    def getVolume(self):
        """TSDF as (value, weight) planes of shape (Z, Y, X) in storage type."""
        n = self.configuration.volumeSize
        dt = self._plane_dtype()
        v = np.zeros((n[2], n[1], n[0]), dtype=dt)
        w = np.zeros((n[2], n[1], n[0]), dtype=dt)
        K.check(K.lib().kf_read_tsdf(self.session, v.ctypes.data, w.ctypes.data))
        return v, w

    def setVolume(self, value, weight):
        dt = self._plane_dtype()
        v = np.ascontiguousarray(value, dtype=dt)
        w = np.ascontiguousarray(weight, dtype=dt)
        K.check(K.lib().kf_write_tsdf(self.session, v.ctypes.data, w.ctypes.data))

    def setDebug(self, on=True):
        K.check(K.lib().kf_set_debug(self.session, 1 if on else 0))

    def readTrackPixels(self, level):
        w, h = self._size[0] >> level, self._size[1] >> level
        res = np.zeros(h * w, dtype=np.int32)
        err = np.zeros(h * w, dtype=np.float32)
        J = np.zeros((h * w, 6), dtype=np.float32)
        K.check(K.lib().kf_read_track_pixels(self.session, level, res.ctypes.data, err.ctypes.data,
                                             J.ctypes.data))
        return res, err, J

    def countIntegrateTraffic(self):
        r, w = C.c_uint64(), C.c_uint64()
        K.check(K.lib().kf_count_integrate_traffic(self.session, C.byref(r), C.byref(w)))
        return r.value, w.value

    def readIcpSolution(self):
        """The 6-vector of the last ICP solve (translation, then angles)."""
        x = np.zeros(6, dtype=np.float64)
        K.check(K.lib().kf_read_icp_solution(self.session, x.ctypes.data))
        return x

    def countIntegrateFastTraffic(self):
        """(voxels read, voxels written) by the default integrate for the
        current depth frame and pose."""
        r, w = C.c_uint64(), C.c_uint64()
        K.check(K.lib().kf_count_integrate_fast_traffic(self.session, C.byref(r), C.byref(w)))
        return r.value, w.value

    def setOption(self, option, value):
        """kf_set_option: K.OPT_FAST_MATH, OPT_SKIP_EMPTY, OPT_GRAPH, OPT_BRICK_REBUILD_PERIOD."""
        K.check(K.lib().kf_set_option(self.session, int(option), int(value)))

    def getOption(self, option):
        v = C.c_int()
        K.check(K.lib().kf_get_option(self.session, int(option), C.byref(v)))
        return v.value

    def raycastStats(self):
        st = (C.c_uint64 * 6)()
        K.check(K.lib().kf_raycast_stats(self.session, st))
        return dict(evaluated=st[0], skipped=st[1], bricks=st[2], hits=st[3], max_eval=st[4], max_iter=st[5])

    def raycastRayCounts(self):
        """Per-pixel (evaluated samples, empty cells entered) of one more
        raycast(), as two (h, w) uint16 arrays."""
        w, h = self._size
        out = np.zeros(w * h, np.uint32)
        K.check(K.lib().kf_raycast_ray_counts(self.session, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        out = out.reshape(h, w)
        return (out & 0xFFFF).astype(np.uint16), (out >> 16).astype(np.uint16)

    def raycastWaveTrace(self):
        """Per-wavefront records of one more raycast(): a structured array with
        start / end (100 MHz GPU clock ticks), HW_ID, XCC id, 16x16 pixel
        tile, lanes per ray (4 for the hybrid schedule's heavy-tile quarters),
        the wavefront's largest loop iteration count, and the block and wave
        numbers; wavefronts that held no pixel are left out."""
        n = C.c_int()
        raw = np.zeros(8 * 4 * 4096, np.uint32)
        K.check(K.lib().kf_raycast_wave_trace(self.session, raw.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(n)))
        r = raw[:8 * n.value].reshape(-1, 8)
        k = np.nonzero((r[:, 7] >> 24) == 0xAB)[0]
        r = r[k].astype(np.uint64)
        out = np.zeros(len(k), dtype=[("start", np.uint64), ("end", np.uint64), ("hw_id", np.uint32),
                                      ("xcc", np.uint32), ("tile", np.int32), ("lanes", np.int32),
                                      ("iters", np.int32), ("block", np.int32), ("wave", np.int32)])
        out["start"] = r[:, 0] | (r[:, 1] << np.uint64(32))
        out["end"] = r[:, 2] | (r[:, 3] << np.uint64(32))
        out["hw_id"], out["xcc"], out["tile"] = r[:, 4], r[:, 5], r[:, 6]
        out["lanes"], out["iters"] = (r[:, 7] >> np.uint64(16)) & np.uint64(0xFF), r[:, 7] & np.uint64(0xFFFF)
        out["block"], out["wave"] = k // 4, k % 4
        return out

    def countRaycastTraffic(self):
        """(distinct voxels read, distinct occupancy-flag bytes read) by
        raycast() for the current volume and pose."""
        v, f = C.c_uint64(), C.c_uint64()
        K.check(K.lib().kf_count_raycast_traffic(self.session, C.byref(v), C.byref(f)))
        return v.value, f.value

    def stageTimes(self):
        """Last pyramid / ICP / integrate / raycast durations in ms (needs
        setOption(OPT_STAGE_TIMING, 1))."""
        ms = (C.c_float * 4)()
        K.check(K.lib().kf_read_stage_times(self.session, ms))
        return dict(pyramid=ms[0], icp=ms[1], integrate=ms[2], raycast=ms[3])

    def rebuildBricks(self):
        K.check(K.lib().kf_rebuild_bricks(self.session))

    def getBrickFlags(self):
        """(occupancy, zero) 8^3 brick flags as (NB2, NB1, NB0) uint8 arrays."""
        nb = [(n + 7) // 8 for n in self.configuration.volumeSize]
        occ = np.zeros(nb[::-1], np.uint8)
        zero = np.zeros(nb[::-1], np.uint8)
        K.check(K.lib().kf_read_brick_flags(self.session, occ.ctypes.data_as(C.POINTER(C.c_uint8)),
                                            zero.ctypes.data_as(C.POINTER(C.c_uint8))))
        return occ, zero

    def synchronize(self):
        K.check(K.lib().kf_synchronize(self.session))
