"""CPU oracle (test infrastructure only): ctypes wrapper of kf_oracle.c.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package, as the checker or the CPU baseline.  The product library never does.

Arrays use the oracle's layouts: maps are (H, W, 4) in R (float32 for the
FP32 oracle, float64 for the FP64 oracle); poses are 4x4 row-major numpy
arrays converted to glm's column-major layout at the boundary; the TSDF is
interleaved (value, weight) pairs, shape (Z, Y, X, 2).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

RG16I, F32X2, F16X2, F64X2 = 0, 1, 2, 3
LAST_ROW, SUM_ALL = 0, 1
_VOL_DTYPES = {RG16I: np.int16, F32X2: np.float32, F16X2: np.uint16, F64X2: np.float64}


def build():
    """Compiles libkforacle_f32.so / libkforacle_f64.so with the Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Oracle:
    """One oracle flavour: precision 'f32' (shader semantics) or 'f64'."""

    def __init__(self, precision="f32"):
        path = os.path.join(HERE, "libkforacle_%s.so" % precision)
        if not os.path.exists(path):
            build()
        self.L = C.CDLL(path)
        self.rs = self.L.kfo_real_size()
        self.dt = np.float32 if self.rs == 4 else np.float64
        self.R = C.c_float if self.rs == 4 else C.c_double
        self._declare()

    # -- ctypes plumbing ---------------------------------------------------
    def _declare(self):
        L, vp, i, R = self.L, C.c_void_p, C.c_int, self.R
        sig = {
            "kfo_f2h": ([C.c_float], C.c_uint16),
            "kfo_h2f": ([C.c_uint16], C.c_float),
            "kfo_camera_matrix": ([vp, vp], None),
            "kfo_inverse_camera_matrix": ([vp, vp], None),
            "kfo_mul4": ([vp, vp, vp], None),
            "kfo_inverse4": ([vp, vp], None),
            "kfo_mip_depth": ([vp, i, i, vp], None),
            "kfo_depth_to_vertex": ([vp, i, i, vp, vp], None),
            "kfo_vertex_to_normal": ([vp, i, i, vp], None),
            "kfo_track_view": ([vp, vp, vp], None),
            "kfo_track": ([vp, vp, i, i, vp, vp, i, i, vp, vp, R, R, vp, vp, vp], None),
            "kfo_reduce": ([vp, vp, vp, i, i, vp], None),
            "kfo_get_reduction": ([vp, i, vp, vp, vp], None),
            "kfo_solve6": ([vp, vp, vp], None),
            "kfo_pose_update": ([vp, vp], None),
            "kfo_voxel_bytes": ([i], i),
            "kfo_reset_volume": ([vp, i, vp], None),
            "kfo_integrate": ([vp, i, vp, vp, vp, i, i, vp, vp, R, R], None),
            "kfo_raycast_view": ([vp, vp, vp], None),
            "kfo_raycast": ([vp, i, vp, vp, vp, vp, R, R, R, R, i, i, vp, vp], None),
            "kfo_config_size": ([], i),
            "kfo_session_create": ([vp], vp),
            "kfo_session_destroy": ([vp], None),
            "kfo_session_set_pose": ([vp, vp], None),
            "kfo_session_get_pose": ([vp, vp], None),
            "kfo_session_buffer": ([vp, i, i], vp),
            "kfo_session_stats": ([vp, vp, vp, vp, vp, vp, vp], None),
            "kfo_session_raycast": ([vp], None),
            "kfo_session_reset": ([vp, vp], None),
            "kfo_session_depth": ([vp, vp], None),
            "kfo_session_track": ([vp], i),
            "kfo_session_integrate": ([vp], None),
            "kfo_session_frame": ([vp, vp], i),
            "kfo_sdf_sample": ([vp, i, vp, R, R, R], R),
            "kfo_track_sdf_pixel": ([vp, i, vp, vp, vp, vp, vp], None),
            "kfo_track_sdf": ([vp, i, vp, vp, vp, i, i, vp, vp], None),
            "kfo_reduce_sdf": ([vp, i, i, vp], None),
            "kfo_ldlt_solve6": ([vp, vp, vp], None),
            "kfo_twist_exp": ([vp, vp], None),
            "kfo_sdf_step": ([vp, vp, i, vp, vp], C.c_double),
            "kfo_session_track_sdf": ([vp], i),
            "kfo_session_recover_pose": ([vp], i),
            "kfo_session_set_mode": ([vp, i, i], None),
            "kfo_session_lib_size": ([vp], i),
            "kfo_intensity_projection": ([vp, i, vp, vp, vp, vp, R, R, R, i, i, vp], None),
            "kfo_mc_tables": ([vp, vp], None),
            "kfo_mc_classify": ([vp, i, vp, R, vp, vp], None),
            "kfo_mc_extract": ([vp, i, i, R, vp, C.c_long], C.c_long),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res

    def arr(self, a, dtype=None):
        return np.ascontiguousarray(a, dtype=dtype or self.dt)

    @staticmethod
    def p(a):
        return a.ctypes.data

    def cm(self, M):
        """row-major 4x4 -> column-major 16 (R)"""
        return self.arr(np.asarray(M).T).reshape(16)

    def rm(self, v):
        return np.asarray(v).reshape(4, 4).T.copy()

    # -- primitives --------------------------------------------------------
    # Every converted array is bound to a local name before the C call: a
    # temporary whose address was taken could be freed (and its memory reused
    # by the next conversion) before the call reads it.
    def f2h(self, x):
        return self.L.kfo_f2h(float(x))

    def h2f(self, h):
        return self.L.kfo_h2f(int(h))

    def camera_matrix(self, cam):
        out, c = np.zeros(16, self.dt), self.arr(cam)
        self.L.kfo_camera_matrix(self.p(c), self.p(out))
        return self.rm(out)

    def inverse_camera_matrix(self, cam):
        out, c = np.zeros(16, self.dt), self.arr(cam)
        self.L.kfo_inverse_camera_matrix(self.p(c), self.p(out))
        return self.rm(out)

    def mul4(self, A, B):
        out, a, b = np.zeros(16, self.dt), self.cm(A), self.cm(B)
        self.L.kfo_mul4(self.p(a), self.p(b), self.p(out))
        return self.rm(out)

    def inverse4(self, M):
        out, m = np.zeros(16, self.dt), self.cm(M)
        self.L.kfo_inverse4(self.p(m), self.p(out))
        return self.rm(out)

    def mip_depth(self, d):
        d = self.arr(d)
        h, w = d.shape
        out = np.zeros((h // 2, w // 2), self.dt)
        self.L.kfo_mip_depth(self.p(d), w, h, self.p(out))
        return out

    def depth_to_vertex(self, d, cam_l):
        d, c = self.arr(d), self.arr(cam_l)
        h, w = d.shape
        out = np.zeros((h, w, 4), self.dt)
        self.L.kfo_depth_to_vertex(self.p(d), w, h, self.p(c), self.p(out))
        return out

    def vertex_to_normal(self, v):
        v = self.arr(v)
        h, w = v.shape[:2]
        out = np.zeros((h, w, 4), self.dt)
        self.L.kfo_vertex_to_normal(self.p(v), w, h, self.p(out))
        return out

    def track_view(self, pose, cam):
        out, P, c = np.zeros(16, self.dt), self.cm(pose), self.arr(cam)
        self.L.kfo_track_view(self.p(P), self.p(c), self.p(out))
        return self.rm(out)

    def track(self, vert, norm, refV, refN, pose, cam, dist_th=0.05, normal_th=0.8):
        vert, norm, refV, refN = (self.arr(a) for a in (vert, norm, refV, refN))
        h, w = vert.shape[:2]
        rh, rw = refV.shape[:2]
        view = self.cm(self.track_view(pose, cam))
        P = self.cm(pose)
        res = np.zeros(h * w, np.int32)
        err = np.zeros(h * w, self.dt)
        J = np.zeros((h * w, 6), self.dt)
        self.L.kfo_track(self.p(vert), self.p(norm), w, h, self.p(refV), self.p(refN), rw, rh,
                         self.p(P), self.p(view), dist_th, normal_th,
                         self.p(res), self.p(err), self.p(J))
        return res, err, J

    def reduce(self, res, err, J, w, h):
        out = np.zeros(256, self.dt)
        r, e, j = self.arr(res, np.int32), self.arr(err), self.arr(J)
        self.L.kfo_reduce(self.p(r), self.p(e), self.p(j), w, h, self.p(out))
        return out.reshape(8, 32)

    def get_reduction(self, partials, rule=LAST_ROW):
        b = np.zeros(6, self.dt)
        Cm = np.zeros(36, self.dt)
        ae = np.zeros(1, self.dt)
        pp = self.arr(partials).reshape(256)
        self.L.kfo_get_reduction(self.p(pp), rule, self.p(b), self.p(Cm), self.p(ae))
        return b, Cm.reshape(6, 6), ae[0]

    def solve6(self, Cm, b):
        x = np.zeros(6, np.float64)
        c, bb = self.arr(Cm, np.float64).reshape(36), self.arr(b, np.float64)
        self.L.kfo_solve6(self.p(c), self.p(bb), self.p(x))
        return x

    def pose_update(self, x, pose):
        P, xx = self.cm(pose), self.arr(x, np.float64)
        self.L.kfo_pose_update(self.p(xx), self.p(P))
        return self.rm(P)

    def voxel_bytes(self, mode):
        return self.L.kfo_voxel_bytes(mode)

    def new_volume(self, N, mode):
        return np.zeros((N[2], N[1], N[0], 2), _VOL_DTYPES[mode])

    def integrate(self, vol, mode, dim, depth_mm, pose, cam, mu, max_weight=100.0):
        assert vol.flags.c_contiguous and vol.dtype == _VOL_DTYPES[mode]
        N = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], np.int32)
        d, dd, P, c = self.arr(depth_mm), self.arr(dim), self.cm(pose), self.arr(cam)
        self.L.kfo_integrate(self.p(vol), mode, self.p(N), self.p(dd), self.p(d), d.shape[1], d.shape[0],
                             self.p(P), self.p(c), mu, max_weight)
        return vol

    def raycast(self, vol, mode, dim, pose, cam, near, far, step, large_step, w=512, h=424):
        N = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], np.int32)
        rv = np.zeros((h, w, 4), self.dt)
        rn = np.zeros((h, w, 4), self.dt)
        v, dd, P, c = np.ascontiguousarray(vol), self.arr(dim), self.cm(pose), self.arr(cam)
        self.L.kfo_raycast(self.p(v), mode, self.p(N), self.p(dd), self.p(P), self.p(c), near, far, step,
                           large_step, w, h, self.p(rv), self.p(rn))
        return rv, rn

    # -- point-to-volume tracking --------------------------------------------
    @staticmethod
    def _N(vol):
        return np.array([vol.shape[2], vol.shape[1], vol.shape[0]], np.int32)

    def sdf_sample(self, vol, mode, loc):
        v, N = np.ascontiguousarray(vol), self._N(vol)
        return self.L.kfo_sdf_sample(self.p(v), mode, self.p(N), float(loc[0]), float(loc[1]), float(loc[2]))

    def track_sdf_pixel(self, vol, mode, dim, T, v):
        vv, N, dd, P, pt = np.ascontiguousarray(vol), self._N(vol), self.arr(dim), self.cm(T), self.arr(v[:3])
        rec = np.zeros(8, self.dt)
        self.L.kfo_track_sdf_pixel(self.p(vv), mode, self.p(N), self.p(dd), self.p(P), self.p(pt), self.p(rec))
        return rec

    def track_sdf(self, vol, mode, dim, vert, T, buf=None):
        """trackSDF over one level: returns the (h*w, 8) record buffer (records
        of the last column are left as they were in `buf`, zero if None)."""
        vv, N, dd, P, vt = np.ascontiguousarray(vol), self._N(vol), self.arr(dim), self.cm(T), self.arr(vert)
        h, w = vt.shape[:2]
        out = np.zeros((h * w, 8), self.dt) if buf is None else self.arr(buf).reshape(h * w, 8).copy()
        self.L.kfo_track_sdf(self.p(vv), mode, self.p(N), self.p(dd), self.p(vt), w, h, self.p(P), self.p(out))
        return out

    def reduce_sdf(self, buf, w, h):
        b = self.arr(buf).reshape(-1)
        out = np.zeros(256, self.dt)
        self.L.kfo_reduce_sdf(self.p(b), w, h, self.p(out))
        return out.reshape(8, 32)

    def ldlt_solve6(self, A, b):
        x = np.zeros(6, np.float64)
        a, bb = self.arr(A, np.float64).reshape(36), self.arr(b, np.float64)
        self.L.kfo_ldlt_solve6(self.p(a), self.p(bb), self.p(x))
        return x

    def twist_exp(self, xi):
        out, x = np.zeros(16, self.dt), self.arr(xi, np.float64)
        self.L.kfo_twist_exp(self.p(x), self.p(out))
        return self.rm(out)

    def sdf_step(self, b, Cm, iteration, result, result_prev):
        """One TrackSDF Gauss-Newton step; returns (cnorm, result, result_prev)."""
        bb, cc = self.arr(b), self.arr(Cm).reshape(36)
        r, rp = self.arr(result, np.float64).copy(), self.arr(result_prev, np.float64).copy()
        cn = self.L.kfo_sdf_step(self.p(bb), self.p(cc), int(iteration), self.p(r), self.p(rp))
        return cn, r, rp

    def intensity_projection(self, vol, mode, dim, pose, cam, near, far, large_step, w=512, h=424):
        N = np.array([vol.shape[2], vol.shape[1], vol.shape[0]], np.int32)
        out = np.zeros((h, w, 4), self.dt)
        v, dd, P, c = np.ascontiguousarray(vol), self.arr(dim), self.cm(pose), self.arr(cam)
        self.L.kfo_intensity_projection(self.p(v), mode, self.p(N), self.p(dd), self.p(P), self.p(c), near, far,
                                        large_step, w, h, self.p(out))
        return out

    # -- surface export (histopyramid marching cubes) ---------------------
    def mc_tables(self):
        """(tri[256, 16] edge indices with 255 after the last triangle, ntri[256])"""
        tri = np.zeros((256, 16), np.uint8)
        nt = np.zeros(256, np.uint8)
        self.L.kfo_mc_tables(self.p(tri), self.p(nt))
        return tri, nt

    def mc_classify(self, vol, mode, iso=0.0):
        v, N = np.ascontiguousarray(vol), self._N(vol)
        nt = np.zeros(vol.shape[:3], np.uint8)
        ci = np.zeros(vol.shape[:3], np.uint8)
        self.L.kfo_mc_classify(self.p(v), mode, self.p(N), float(iso), self.p(nt), self.p(ci))
        return nt, ci

    def mc_extract(self, vol, mode, iso=0.0):
        """Triangle vertices (T, 3, 4) in voxel-index units, (x, y, z, 0)."""
        v = np.ascontiguousarray(vol)
        n = vol.shape[0]
        assert vol.shape[:3] == (n, n, n) and n & (n - 1) == 0
        total = self.L.kfo_mc_extract(self.p(v), mode, n, float(iso), None, 0)
        out = np.zeros((max(total, 1), 3, 4), self.dt)
        self.L.kfo_mc_extract(self.p(v), mode, n, float(iso), self.p(out), total)
        return out[:total]


def _config_struct(R):
    class Cfg(C.Structure):
        _fields_ = [
            ("vol", C.c_int * 3), ("dim", R * 3), ("mu", R), ("maxWeight", R), ("nearPlane", R),
            ("farPlane", R), ("dist_threshold", R), ("normal_threshold", R), ("iterations", C.c_int * 3),
            ("cam", R * 4), ("width", C.c_int), ("height", C.c_int), ("mode", C.c_int), ("rule", C.c_int),
        ]
    return Cfg


class OracleSession:
    """The oracle's gFusion state + the kinect.cpp frame block."""

    def __init__(self, oracle, volume, dim, mu, mode, iterations=(4, 5, 10), cam=(365.5, 365.5, 255.5, 211.5),
                 width=512, height=424, rule=LAST_ROW, max_weight=100.0, near=0.1, far=3.0, dist_th=0.05,
                 normal_th=0.8):
        self.o = oracle
        Cfg = _config_struct(oracle.R)
        assert C.sizeof(Cfg) == oracle.L.kfo_config_size()
        c = Cfg()
        n = volume if isinstance(volume, (tuple, list)) else (volume,) * 3
        for k in range(3):
            c.vol[k] = n[k]
            c.dim[k] = dim
            c.iterations[k] = iterations[k]
        c.mu, c.maxWeight, c.nearPlane, c.farPlane = mu, max_weight, near, far
        c.dist_threshold, c.normal_threshold = dist_th, normal_th
        for k in range(4):
            c.cam[k] = cam[k]
        c.width, c.height, c.mode, c.rule = width, height, mode, rule
        self.cfg = c
        self.N = n
        self.w, self.h = width, height
        self.mode = mode
        self.s = oracle.L.kfo_session_create(C.byref(c))

    def __del__(self):
        try:
            self.o.L.kfo_session_destroy(self.s)
        except Exception:
            pass

    def _buf(self, which, level, shape, dtype):
        ptr = self.o.L.kfo_session_buffer(self.s, which, level)
        n = int(np.prod(shape))
        buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype=dtype).reshape(shape)

    def reset(self, pose):
        P = self.o.cm(pose)
        self.o.L.kfo_session_reset(self.s, self.o.p(P))

    def set_pose(self, pose):
        P = self.o.cm(pose)
        self.o.L.kfo_session_set_pose(self.s, self.o.p(P))

    def pose(self):
        out = np.zeros(16, self.o.dt)
        self.o.L.kfo_session_get_pose(self.s, self.o.p(out))
        return self.o.rm(out)

    def depth(self, depth_mm):
        d = np.ascontiguousarray(depth_mm, dtype=np.float32)
        self.o.L.kfo_session_depth(self.s, d.ctypes.data)

    def track(self):
        return self.o.L.kfo_session_track(self.s)

    def integrate(self):
        self.o.L.kfo_session_integrate(self.s)

    def raycast(self):
        self.o.L.kfo_session_raycast(self.s)

    def frame(self, depth_mm):
        d = np.ascontiguousarray(depth_mm, dtype=np.float32)
        return self.o.L.kfo_session_frame(self.s, d.ctypes.data)

    def set_mode(self, track_mode=1, recover=1):
        """track_mode: 1 = Track (point-to-plane), 2 = TrackSDF, 3 = both;
        recover: call recoverPose() on a frame whose tracking failed."""
        self.o.L.kfo_session_set_mode(self.s, int(track_mode), int(recover))

    def track_sdf(self):
        return self.o.L.kfo_session_track_sdf(self.s)

    def recover_pose(self):
        return self.o.L.kfo_session_recover_pose(self.s)

    def library(self):
        n = self.o.L.kfo_session_lib_size(self.s)
        lib = self._buf(14, 0, (256, 16), self.o.dt)[:n]
        return np.array([self.o.rm(p) for p in lib]).reshape(n, 4, 4)

    def pose_eig(self):
        return self.o.rm(self._buf(15, 0, (16,), self.o.dt).copy())

    def set_pose_eig(self, pose):
        self._buf(15, 0, (16,), self.o.dt)[...] = self.o.cm(pose)

    def cum_twist(self):
        return self._buf(16, 0, (6,), np.float64).copy()

    def sdf_buffer(self):
        return self._buf(11, 0, (self.h * self.w, 8), self.o.dt).copy()

    def set_sdf_buffer(self, buf):
        self._buf(11, 0, (self.h * self.w, 8), self.o.dt)[...] = buf

    def sdf_partials(self):
        return self._buf(12, 0, (8, 32), self.o.dt).copy()

    def sdf_log(self):
        return self._buf(13, 0, (64, 10), np.float64).copy()

    def stats(self):
        ae = np.zeros(1, self.o.dt)
        ints = [C.c_int() for _ in range(5)]
        self.o.L.kfo_session_stats(self.s, ae.ctypes.data, *[C.addressof(v) for v in ints])
        t, it, ln, cnt, rst = (v.value for v in ints)
        return dict(ae=float(ae[0]), tracked=t, iterations=it, log_n=ln, counter=cnt, reset=rst)

    def log(self):
        n = self.stats()["log_n"]
        return self._buf(10, 0, (64, 8), np.float64)[:n].copy()

    def map(self, which, level=0):
        w, h = self.w >> level, self.h >> level
        if which == 0:
            return self._buf(0, level, (h, w), self.o.dt).copy()
        return self._buf(which, level, (h, w, 4), self.o.dt).copy()

    def set_map(self, which, arr, level=0):
        w, h = self.w >> level, self.h >> level
        dst = self._buf(which, level, (h, w, 4), self.o.dt)
        dst[...] = arr

    def volume(self):
        n = self.N
        return self._buf(5, 0, (n[2], n[1], n[0], 2), _VOL_DTYPES[self.mode]).copy()

    def set_volume(self, vol):
        n = self.N
        dst = self._buf(5, 0, (n[2], n[1], n[0], 2), _VOL_DTYPES[self.mode])
        dst[...] = vol

    def partials(self):
        return self._buf(9, 0, (8, 32), self.o.dt).copy()

    def track_pixels(self, level):
        w, h = self.w >> level, self.h >> level
        return (self._buf(6, 0, (h * w,), np.int32).copy(), self._buf(7, 0, (h * w,), self.o.dt).copy(),
                self._buf(8, 0, (h * w, 6), self.o.dt).copy())
