"""ctypes binding of libkfusion.so (the C ABI of include/kfusion.h).

The library is the product path: there is no CPU fallback.  Loading fails
loudly if the HIP library has not been built.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libkfusion.so")

KF_OK = 0
KF_ERR_INVALID = -1
KF_ERR_HIP = -2
KF_ERR_OOM = -3
KF_ERR_BUDGET = -4
KF_ERR_STATE = -5

STORAGE_RG16I = 0
STORAGE_F32 = 1
STORAGE_F16 = 2
STORAGE_NAMES = {"rg16i": STORAGE_RG16I, "f32": STORAGE_F32, "f16": STORAGE_F16}

RULE_LAST_ROW = 0
RULE_SUM_ALL = 1

MAP_DEPTH = 0
MAP_VERTEX = 1
MAP_NORMAL = 2
MAP_REF_VERTEX = 3
MAP_REF_NORMAL = 4
MAP_INTENSITY = 5

INTEGRATE_ALWAYS = 0
INTEGRATE_GATED = 1

OPT_FAST_MATH = 1
OPT_SKIP_EMPTY = 2
OPT_GRAPH = 3
OPT_BRICK_REBUILD_PERIOD = 4
OPT_RAYCAST_XCD_GROUP = 5
OPT_ICP_ROWS_PER_BLOCK = 6
OPT_STAGE_TIMING = 7
OPT_TRACK_MODE = 8
OPT_RECOVER = 9
OPT_ICP_ALL_GROUPS = 10
OPT_RAYCAST_LANES = 11
OPT_ICP_PERSISTENT = 12
OPT_ICP_SOLO_ROWS = 13
OPT_ICP_REFERENCE_ORDER = 16
OPT_INTEGRATE_ZLEN = 17
OPT_INTEGRATE_ROWS = 18
OPT_POSE_FP64 = 19
OPT_INTEGRATE_COLUMNS = 20
OPT_RAYCAST_HYBRID = 21
OPT_RAYCAST_HEAVY_TILES = 22

TRACK_POINT_TO_PLANE = 1
TRACK_POINT_TO_VOLUME = 2

POSE_LIBRARY_MAX = 256


class KfConfig(C.Structure):
    """kf_config of include/kfusion.h (gFusionConfig + camera + depth size)."""

    _fields_ = [
        ("volume_size", C.c_int * 3),
        ("volume_dim", C.c_float * 3),
        ("mu", C.c_float),
        ("max_weight", C.c_float),
        ("near_plane", C.c_float),
        ("far_plane", C.c_float),
        ("dist_threshold", C.c_float),
        ("normal_threshold", C.c_float),
        ("iterations", C.c_int * 3),
        ("cam", C.c_float * 4),
        ("width", C.c_int),
        ("height", C.c_int),
        ("storage", C.c_int),
        ("reduce_rule", C.c_int),
    ]


class KfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("kfusion error %d: %s" % (code, msg))
        self.code = code


_lib = None


def lib():
    """Loads libkfusion.so (once) and declares the entry points."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libkfusion.so is not built (%s); run __graft_entry__.build() first" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, i32, f32p, sz = C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_size_t
    sess = C.c_void_p
    sigs = {
        "kf_config_default": [C.POINTER(KfConfig)],
        "kf_session_bytes": [C.POINTER(KfConfig), C.POINTER(C.c_size_t)],
        "kf_session_create": [C.POINTER(KfConfig), vp, C.POINTER(C.c_void_p)],
        "kf_session_destroy": [sess],
        "kf_get_config": [sess, C.POINTER(KfConfig)],
        "kf_reset": [sess, f32p],
        "kf_reset_volume": [sess],
        "kf_set_pose": [sess, f32p],
        "kf_get_pose": [sess, f32p],
        "kf_depth_to_vertex_f32": [sess, vp],
        "kf_depth_to_vertex_u16": [sess, vp],
        "kf_depth_to_vertex_u16_dev": [sess, vp],
        "kf_depth_to_vertex_f32_dev": [sess, vp],
        "kf_track": [sess, C.POINTER(C.c_int)],
        "kf_track_frame": [sess],
        "kf_icp_iteration": [sess, i32, i32, i32],
        "kf_icp_finish": [sess],
        "kf_integrate": [sess, i32],
        "kf_raycast": [sess],
        "kf_frame_u16_dev": [sess, vp],
        "kf_frame_f32": [sess, vp],
        "kf_read_map": [sess, i32, i32, vp, sz],
        "kf_write_map": [sess, i32, i32, vp, sz],
        "kf_read_tsdf": [sess, vp, vp],
        "kf_write_tsdf": [sess, vp, vp],
        "kf_read_reduction": [sess, f32p, f32p],
        "kf_read_status": [sess, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                           C.POINTER(C.c_int)],
        "kf_set_debug": [sess, i32],
        "kf_read_track_pixels": [sess, i32, vp, vp, vp],
        "kf_synchronize": [sess],
        "kf_get_stream": [sess, C.POINTER(C.c_void_p)],
        "kf_count_integrate_traffic": [sess, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
        "kf_count_integrate_fast_traffic": [sess, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
        "kf_read_icp_solution": [sess, C.c_void_p],
        "kf_set_option": [sess, i32, i32],
        "kf_get_option": [sess, i32, C.POINTER(C.c_int)],
        "kf_rebuild_bricks": [sess],
        "kf_read_brick_flags": [sess, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)],
        "kf_read_stage_times": [sess, C.POINTER(C.c_float)],
        "kf_raycast_stats": [sess, C.POINTER(C.c_uint64)],
        "kf_raycast_ray_counts": [sess, C.POINTER(C.c_uint32)],
        "kf_raycast_wave_trace": [sess, C.POINTER(C.c_uint32), C.POINTER(C.c_int)],
        "kf_count_raycast_traffic": [sess, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
        "kf_track_sdf": [sess, C.POINTER(C.c_int)],
        "kf_recover_pose": [sess, C.POINTER(C.c_int)],
        "kf_read_pose_library": [sess, f32p, i32, C.POINTER(C.c_int)],
        "kf_get_cum_twist": [sess, f32p],
        "kf_write_pose_eig": [sess, f32p],
        "kf_mesh_generate": [sess, C.c_float, C.POINTER(C.c_uint64)],
        "kf_intensity_projection": [sess],
        "kf_mesh_read": [sess, f32p, C.c_uint64],
        "kf_mesh_generate_noalloc": [sess, C.c_float, C.c_void_p, C.c_uint64, C.c_void_p],
        "kf_mesh_export_stl": [sess, C.c_char_p],
        "kf_mesh_export_ply": [sess, C.c_char_p],
        "kf_read_sdf_state": [sess, f32p, f32p, f32p, f32p, C.POINTER(C.c_int), C.POINTER(C.c_double)],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.kf_load_calibration.argtypes = [C.c_char_p, vp]
    L.kf_load_calibration.restype = C.c_int
    L.kf_mc_tables.argtypes = [vp, vp]
    L.kf_mc_tables.restype = C.c_int
    L.kf_last_error.argtypes = []
    L.kf_last_error.restype = C.c_char_p
    L.kf_version.argtypes = []
    L.kf_version.restype = C.c_char_p
    batch = {
        "kf_batch_create": [C.POINTER(C.c_void_p), i32, vp, C.POINTER(C.c_void_p)],
        "kf_batch_set_sessions": [vp, C.POINTER(C.c_void_p), i32],
        "kf_batch_destroy": [vp],
        "kf_batch_synchronize": [vp],
        "kf_batch_get_stream": [vp, C.POINTER(C.c_void_p)],
        "kf_batch_frame_u16_dev": [vp, C.POINTER(C.c_void_p), C.POINTER(C.c_int)],
        "kf_batch_set_option": [vp, i32, i32],
        "kf_batch_get_option": [vp, i32, C.POINTER(C.c_int)],
    }
    for name, args in batch.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != KF_OK:
        msg = lib().kf_last_error()
        raise KfError(rc, msg.decode() if msg else "")
    return rc


def exported_symbols():
    """Names declared in include/kfusion.h that the library must export."""
    import re
    hdr = os.path.join(os.path.dirname(PKG), "include", "kfusion.h")
    names = []
    with open(hdr) as f:
        for line in f:
            m = re.match(r"^\s*(?:const\s+)?(?:int|const char\s*\*|char\s*\*)\s*\*?\s*(kf_\w+)\s*\(", line)
            if m:
                names.append(m.group(1))
    return names
