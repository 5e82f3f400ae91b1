// kf_internal.hpp -- kernel argument blocks and launchers shared by the
// kernels (kf_kernels.hip) and the session / C-ABI layer (kf_capi.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/kfusion.h"

namespace kf {

struct DevState;

// The hybrid raycast's schedule of a session: per 16x16 tile the largest
// one-lane loop iteration count since the last selection, the tiles selected
// for four lanes per ray (heavy[0] = count, heavy[1..] = tile numbers) and
// their flags.  tiles = 0: no schedule.
struct RaySched {
    uint32_t *tile_cost;
    int *heavy;
    uint8_t *heavy_flag;
    int tiles;
    int heavy_max;
};

struct PyramidArgs {
    int w, h;               // level-0 size
    float cam[4];           // level-0 fx, fy, cx, cy
    int block_start[3];     // first block of each level in the 1-D grid
    int blocks_total;
    float *depth0;          // level-0 depth (float mm)
    float *depth_m;         // level-0 depth / 1000 (metres), read by integrate
    float4 *vert[3];
    float4 *norm[3];
    float2 *dtiles;         // per 8x8 level-0 tile: (min, max) of its nonzero depth_m
    int tiles_x, tiles_y;
    // k_pyramid with sched.tiles > 0: block 0 selects the hybrid raycast's
    // next heavy tiles (the pyramid's blocks start at 1)
    RaySched sched;
};

struct IcpArgs {
    const float4 *vert, *norm;  // level maps
    int w, h;                   // level size
    const float4 *refV, *refN;  // level-0 raycast maps
    int rw, rh;
    float cam[4];
    float dist_th, normal_th;
    int first, save_pose, rule;
    int finish, frame;          // last launch of Track(): accept / reject (+ frame gate) at its end
    float *block_partials;      // [8 * rows_step][32] (28 used)
    float *library;             // pose library [lib_max][16] (updatePoseFinder on success)
    int lib_max;
    int rows_step;              // blocks per row group
    int group_only;             // -1: all 8 row groups; g: only row group g (the one the combine rule reads)
    int pose_fp64;              // compose the pose update in FP64 (KF_OPT_POSE_FP64)
    int fast_trig;              // angles' cos / sin by series for |x| <= 0.5 (off in the reference-order mode)
    int32_t *dbg_result;        // optional per-pixel output
    float *dbg_error, *dbg_J;
};

// Track() in one persistent launch (k_icp_track): every level's arguments,
// its iteration count, and two partial-row buffers used alternately.
struct IcpTrackArgs {
    IcpArgs lv[3];
    int iters[3];
    int frame;                  // apply the frame gate at the end
    unsigned long long *rows[2];  // partial rows, alternately: per term (tag << 32) | float bits
    unsigned spin_limit;        // grid-barrier / row wait bound (polls)
};

struct IntegrateArgs {
    int N[3];
    float dim[3];
    float cam[4];
    float mu, max_weight;
    const float *depth_m; // level-0 depth / 1000 (metres)
    int dw, dh;
    void *tsdf, *weight;  // planes
    int zlen;             // z voxels per segment (blockIdx.z), multiple of 8: voxels per thread
    int rows;             // y rows per block (blockDim.y: 1, 2 or 4), 64 x per row
    int gated;
    uint8_t *bflags;      // 8^3 brick occupancy flags (set-only)
    uint8_t *zflags;      // 8^3 bricks that may hold a voxel other than (0, 0) (set-only)
    int NB[3];            // brick grid
    const float2 *dtiles; // (min, max) nonzero depth_m per 8x8 pixel tile
    int tiles_x, tiles_y;
    // per-column set-up computed once per frame by k_integrate_columns
    // (nullptr: each thread computes its column's): the column's camera-frame
    // position at z = 0, its frustum range before the segment clip, and the
    // z step (the same for every column)
    float4 *col_pos;
    int2 *col_range;
    float *col_dl;
};

struct BrickArgs {
    int N[3], NB[3], NC[3];
    const void *tsdf, *weight;
    uint8_t *flags, *dflags, *cflags, *zflags;
    // k_brick_rebuild: 1 = only bricks whose zero flag is set are read (the
    // flags are exact supersets since the last rebuild); 0 = every brick
    int incremental;
};

struct RaycastArgs {
    int N[3];
    float dim[3];
    float cam[4];
    float near_plane, far_plane, step, large_step;
    const void *tsdf;
    int w, h;
    int tiles_x, tiles_y, blocks;
    float4 *refV, *refN;
    const uint8_t *dflags;  // dilated brick flags
    const uint8_t *cflags;  // 32^3 coarse cells (OR of dflags)
    int NB[3], NC[3];
    int skip;               // empty-space skipping on / off
    unsigned long long *stats;  // optional: evaluated, skipped, cells skipped, hits, max eval, max iters
    int dim_pow2;               // all volume_dim are powers of two
    float vscale[3];            // N / dim, exact when dim_pow2 (a power-of-two scaling)
    int off32;                  // value-plane bytes < 2^32 and N0 N1 < 2^24 (32-bit offsets, 24-bit products)
    int xcd_group;              // tiles per XCD cluster (block -> tile mapping)
    int lanes;                  // lanes per ray: 1 (16x16-pixel tiles) or 4 (8x8-pixel tiles)
    uint32_t *ray_counts;       // optional (with stats): per pixel, evaluated samples | skip events << 16
    uint32_t *touch;            // k_raycast_touch: bitmap of voxels, then flag bytes, read
    uint32_t *wave_trace;       // optional: 8 words per wavefront (clock at start and end, HW_ID, XCC id, tile)
    // hybrid schedule (k_raycast_hybrid; heavy_max 0: off), as RaySched:
    // the launch adds its tiles' counts and marches the listed tiles with
    // four lanes per ray
    uint32_t *tile_cost;
    int *heavy;
    uint8_t *heavy_flag;
    int heavy_max;              // tiles marched four lanes per ray at most (a multiple of 8)
};

constexpr int kRayHeavyMax = 48;   // hybrid raycast: tiles per launch marched four lanes per ray
constexpr int kRayHeavyFloor = 16; // ... only tiles whose last march took at least this many iterations
// block b < 4 kRayHeavyMax of a hybrid launch takes quarter (b / 8) % 4 of
// heavy tile (b / 32) 8 + b % 8: every heavy tile once when this holds
static_assert(kRayHeavyMax % 8 == 0, "heavy tiles come in groups of 8 (one per XCD)");
// bytes of a session's hybrid-schedule state for `tiles` 16x16 tiles
// (tile_cost[tiles], heavy[1 + kRayHeavyMax], heavy_flag[tiles])
inline size_t ray_sched_bytes(int tiles) { return 4 * ((size_t)tiles + kRayHeavyMax + 1) + (size_t)tiles; }

// Point-to-volume tracking (persistent k_sdf_track, one launch per TrackSDF /
// recoverPose).
constexpr int kSdfTrack = 0;          // TrackSDF() from m_pose_eig
constexpr int kSdfRecover = 1;        // recoverPose(): TrackSDF from each library entry
constexpr int kSdfRecoverIfLost = 2;  // recoverPose() only if the frame's tracking failed
constexpr int kPoseLibraryMax = 256;  // pose library capacity (entries)

struct SdfArgs {
    const float4 *vert[3];  // level vertex maps (camera frame)
    int w[3], h[3];         // level sizes
    int iters[3];           // iterations per level, 0 = finest
    const void *tsdf;       // value plane
    int N[3];
    float dim[3];
    int rule;               // combine rule of getSDFReduction
    int mode;               // kSdf*
    int frame;              // kSdfTrack inside the frame loop: apply the frame gate
    float *partials;        // 2 x nblk rows of 32 floats (double-buffered per iteration)
    float *slots;           // records left at the last-column buffer indices: [H0 + H1 + H2][8]
    int slot_base[3];       // first slot of each level
    const float *library;   // pose library [kPoseLibraryMax][16]
    unsigned spin_limit;    // grid-barrier wait bound (polls)
    int dilate;             // 1: first the dilated brick flags and coarse cells of `bricks` (k_brick_dilate's work)
    BrickArgs bricks;
};
hipError_t launch_sdf_track(const SdfArgs &a, int storage, DevState *st, int nblk, hipStream_t s);

// ---------------------------------------------------------------------------
// Multi-session batches: one launch per kernel type for all active sessions.
// ---------------------------------------------------------------------------
constexpr int kMaxBatch = 64;

// everything one session's kernels need, resident in device memory
struct BatchSession {
    PyramidArgs pyr;
    IcpArgs icp[3];
    IcpTrackArgs trk;  // the persistent Track() of this session (k_batch_icp_track)
    IntegrateArgs integ;
    RaycastArgs ray;
    BrickArgs brick;
    DevState *st;
    int int_grid[3];  // integrate grid of this session
};

// per-tick record, read by the batch kernels from device memory: the depth
// inputs and the active mask
struct BatchTick {
    const uint16_t *depth[kMaxBatch];
    unsigned long long active;
};

// the batch's member tables (kernel arguments; fixed between membership changes)
struct BatchTable {
    const BatchSession *sess;            // n entries
    const int *int_prefix, *coa_prefix;  // n + 1 entries each
    int n;
};

void integrate_grid_dims(const IntegrateArgs &a, int g[3]);
// every byte offset into a TSDF plane of this volume fits in 32 bits: the
// fast integrate then addresses the planes through buffer resources
inline bool integrate_planes_fit_32(const IntegrateArgs &a, int storage)
{
    const size_t vb = storage == KF_STORAGE_F32 ? 4 : 2;
    return (size_t)a.N[0] * a.N[1] * a.N[2] * vb <= (size_t)0xFFFFFFFFu;
}
// tb: the member tables; tp: the tick record in device memory
hipError_t launch_batch_pyramid(const BatchTable &tb, const BatchTick *tp, int blocks_per_session, hipStream_t s);
hipError_t launch_batch_icp(const BatchTable &tb, const BatchTick *tp, int level, int first, int save_pose,
                            int finish, int nblk, hipStream_t s);
hipError_t launch_batch_icp_track(const BatchTable &tb, const BatchTick *tp, int blocks_per_session, hipStream_t s);
hipError_t launch_batch_integrate(const BatchTable &tb, const BatchTick *tp, int storage, int fast, int off32,
                                  int total_blocks, int rows, hipStream_t s);
hipError_t launch_batch_bricks(const BatchTable &tb, const BatchTick *tp, int total_coarse, hipStream_t s);
hipError_t launch_batch_raycast(const BatchTable &tb, const BatchTick *tp, int storage, int lanes,
                                int blocks_per_session, int dim_pow2, hipStream_t s);

// Optional t0 / t1: events recorded by the kernel's own dispatch packet at its
// start / end (exact kernel duration, no separate marker packets).
hipError_t launch_pyramid_u16(const PyramidArgs &a, const uint16_t *in, hipStream_t s, hipEvent_t t0 = nullptr,
                              hipEvent_t t1 = nullptr);
hipError_t launch_pyramid_f32(const PyramidArgs &a, const float *in, hipStream_t s, hipEvent_t t0 = nullptr,
                              hipEvent_t t1 = nullptr);
hipError_t launch_icp_iteration(const IcpArgs &a, DevState *st, hipStream_t s, hipEvent_t t0 = nullptr,
                                hipEvent_t t1 = nullptr);
// blocks of a k_icp_iteration launch for these arguments
__host__ __device__ inline int icp_blocks(const IcpArgs &a) { return (a.group_only >= 0 ? 1 : 8) * a.rows_step; }
// one launch of nblk blocks (nblk >= icp_blocks of every level with
// iterations, all resident at once)
// one ICP iteration with the reference's summation order: per-pixel records
// into a.dbg_* (which must be set), then reduce.cs's sums, combine, solve
hipError_t launch_icp_ref_iteration(const IcpArgs &a, DevState *st, hipStream_t s, hipEvent_t t0 = nullptr,
                                    hipEvent_t t1 = nullptr);
hipError_t launch_icp_track(const IcpTrackArgs &a, DevState *st, int nblk, hipStream_t s, hipEvent_t t0,
                            hipEvent_t t1);
hipError_t launch_icp_finish(DevState *st, int frame, float *lib, int lib_max, hipStream_t s,
                             hipEvent_t t1 = nullptr);
// state flags of launch_state_set (kStateSetPose is implied by a non-NULL pose)
constexpr int kStateSetPose = 1, kStateArmWindow = 2, kStateResetPose = 4;
hipError_t launch_state_set(DevState *st, const float *pose, int flags, hipStream_t s);
hipError_t launch_integrate(IntegrateArgs a, int storage, int fast, const DevState *st, hipStream_t s,
                            hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
hipError_t launch_brick_rebuild(const BrickArgs &a, int storage, hipStream_t s);
hipError_t launch_brick_dilate(const BrickArgs &a, hipStream_t s);
// the voxels k_integrate_fast reads and writes (counts[0], counts[1]), no stores
hipError_t launch_count_integrate_fast(const IntegrateArgs &a, int storage, const DevState *st,
                                       unsigned long long *counts, hipStream_t s);
hipError_t launch_count_integrate(const IntegrateArgs &a, const DevState *st, unsigned long long *counts,
                                  hipStream_t s);
hipError_t launch_intensity_projection(const RaycastArgs &a, int storage, const DevState *st, float4 *out,
                                       hipStream_t s);
// k_raycast_touch then the population count of its `words`-word bitmap into *count
hipError_t launch_raycast_touch(const RaycastArgs &a, int storage, const DevState *st, size_t words,
                                unsigned long long *count, hipStream_t s);
hipError_t launch_raycast(const RaycastArgs &a, int storage, const DevState *st, hipStream_t s,
                          hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);

}  // namespace kf
