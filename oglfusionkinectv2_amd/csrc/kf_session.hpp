// kf_session.hpp -- the session object behind the opaque kf_session handle:
// the device buffers gFusion creates in initTextures / initVolume /
// allocateBuffers (gFusion.cpp:200-330) plus this library's own state.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/kfusion.h"
#include "kf_device.hpp"
#include "kf_internal.hpp"

struct kf_session {
    kf_config cfg;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = 0;
    float *depth0 = nullptr;
    float *depth_m = nullptr;
    uint16_t *stage_u16 = nullptr;
    float *stage_f32 = nullptr;
    // pinned host staging of the host-buffer uploads (two slots, allocated on
    // first use): the caller's frame is copied into a slot and sent by DMA in
    // stream order, so the call returns without waiting for the GPU
    void *h_stage[2] = {nullptr, nullptr};
    hipEvent_t h_stage_ev[2] = {nullptr, nullptr};
    bool h_stage_pending[2] = {false, false};
    int h_stage_next = 0;
    float4 *vert[3] = {nullptr, nullptr, nullptr};
    float4 *norm[3] = {nullptr, nullptr, nullptr};
    float4 *refV = nullptr, *refN = nullptr;
    void *tsdf = nullptr, *weight = nullptr;
    kf::DevState *st = nullptr;
    float *block_partials = nullptr;
    unsigned long long *icp_rows = nullptr;  // tagged partial rows of the persistent Track(): 2 x (height + 8) x 32
    unsigned long long *counts = nullptr;
    uint8_t *bflags = nullptr, *dflags = nullptr, *cflags = nullptr, *zflags = nullptr;
    float2 *dtiles = nullptr;  // 8x8-pixel (min, max) of the level-0 depth in metres
    float4 *col_pos = nullptr;  // integrate's per-column set-up (KF_OPT_INTEGRATE_COLUMNS)
    int2 *col_range = nullptr;
    float *col_dl = nullptr;
    unsigned char *ray_sched = nullptr;  // hybrid raycast schedule: tile costs, heavy list, flags, block count
    int ray_hybrid = 1;                  // KF_OPT_RAYCAST_HYBRID
    // stage timing (KF_OPT_STAGE_TIMING): start / stop events of pyramid, ICP
    // (first iteration .. finish), integrate, raycast, carried by the dispatch packets
    int timing = 0;
    hipEvent_t tev[8] = {};
    int NB[3] = {0, 0, 0}, NC[3] = {0, 0, 0};
    size_t nbricks = 0, ncoarse = 0;
    // options (kf_set_option)
    int fast_math = 1;
    int skip_empty = 1;
    int use_graph = 1;
    int rebuild_period = 16;
    int xcd_group = 4;
    int icp_rows_per_block = 1;
    int ray_lanes = 1;       // raycast lanes per ray (1 or 4)
    int icp_all_groups = 0;  // ICP sums every row group, not only the one the combine rule reads
    int icp_persistent = 1;  // Track() as one persistent launch
    int icp_solo_rows = 0;   // levels with at most this many rows per row group: one block per group
    int icp_ref_order = 0;   // Track() sums in reduce.cs's order (per-pixel records + serial FP32 sums)
    int int_zlen = 0;        // integrate z voxels per thread (0: about 2M column segments per launch)
    int int_rows = 4;        // integrate block rows (64 x int_rows threads)
    int pose_fp64 = 0;       // ICP pose updates composed in FP64 (KF_OPT_POSE_FP64)
    int int_columns = 1;     // per-column integrate set-up once per frame (KF_OPT_INTEGRATE_COLUMNS)
    int cus = 0;             // compute units of the device
    // frame graph: ICP + gate + integrate + raycast, captured once
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_failed = 0;
    long frames = 0;
    int debug = 0;
    int dbg_level = -1;
    int32_t *dbg_result = nullptr;
    float *dbg_error = nullptr, *dbg_J = nullptr;
    size_t voxels = 0;
    // point-to-volume tracking and pose recovery
    float *library = nullptr;       // pose library [kPoseLibraryMax][16]
    float *sdf_partials = nullptr;  // 2 x sdf_blocks rows of 32 floats
    float *sdf_slots = nullptr;     // last-column records, (H0 + H1 + H2) x 8 floats
    int sdf_blocks = 0;             // persistent grid: at most one block per CU, multiple of 8
    int track_mode = 1;             // KF_TRACK_POINT_TO_PLANE / _VOLUME / both
    int recover = 1;                // recoverPose() in the frame loop when tracking fails
    // surface export (MCubes, mcubes.cpp)
    uint8_t *mc_tab = nullptr;          // triangle table [256][16] + triangle counts [256]
    uint32_t *mc_count = nullptr;       // triangles per 8^3 brick, Morton order
    uint32_t *mc_offset = nullptr;      // their exclusive prefix sums
    uint32_t *mc_part = nullptr;        // scan partials
    unsigned long long *mc_total = nullptr;
    size_t mc_slots = 0;                // Morton brick slots allocated
    float4 *mc_verts = nullptr;         // 3 (x, y, z, 0) vertices per triangle, voxel-index units
    size_t mc_cap = 0;                  // triangles mc_verts holds
    unsigned long long mc_triangles = 0;  // triangles of the last kf_mesh_generate
    float4 *mip = nullptr;              // intensityProjection image (RGBA32F, W x H)
    uint8_t *mip_zflags = nullptr;      // its dilated zero flags (nbricks) + coarse cells (ncoarse)
};

// sets the calling thread's error message (kf_last_error); returns code
int kf_fail(int code, const char *msg);

namespace kf {
// once per device (kf_session_create): the device's surface-export scratch
// address and its triangle table (kf_mcubes.hip)
int mesh_noalloc_setup(int device);
}  // namespace kf
