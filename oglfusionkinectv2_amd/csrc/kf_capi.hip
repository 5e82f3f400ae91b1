// kf_capi.hip -- session management and the C ABI declared in include/kfusion.h.
//
// A session owns the device buffers that gFusion creates in initTextures /
// initVolume / allocateBuffers (gFusion.cpp:200-330):
//   depth0            float  W*H            (R32F depth level 0, mm)
//   vert[l], norm[l]  float4 (W>>l)*(H>>l)  (RGBA32F vertex / normal pyramids)
//   refV, refN        float4 W*H            (raycast reference maps)
//   tsdf, weight      planes of N0*N1*N2 voxels, x fastest (RG16I volume)
//   state             DevState (pose, ICP partials, frame flags)
// All work is queued on the session's stream; nothing synchronises except the
// read-back calls and kf_track with a non-NULL result pointer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "kf_device.hpp"
#include "kf_internal.hpp"
#include "kf_session.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define KF_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(KF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

size_t voxel_plane_bytes(int storage)
{
    switch (storage) {
    case KF_STORAGE_RG16I: return 2;
    case KF_STORAGE_F32: return 4;
    case KF_STORAGE_F16: return 2;
    }
    return 0;
}

int validate(const kf_config *c)
{
    if (!c) return fail(KF_ERR_INVALID, "config is NULL");
    for (int k = 0; k < 3; ++k) {
        if (c->volume_size[k] < 1 || c->volume_size[k] > 2048)
            return fail(KF_ERR_INVALID, "volume_size must be in [1, 2048]");
        if (!(c->volume_dim[k] > 0.0f) || !std::isfinite(c->volume_dim[k]))
            return fail(KF_ERR_INVALID, "volume_dim must be positive");
        if (c->iterations[k] < 0 || c->iterations[k] > 64)
            return fail(KF_ERR_INVALID, "iterations must be in [0, 64]");
    }
    if (c->width < 8 || c->height < 8 || c->width % 4 || c->height % 4 || c->width > 8192 || c->height > 8192)
        return fail(KF_ERR_INVALID, "depth size must be multiples of 4 in [8, 8192]");
    if (!(c->mu > 0.0f) || !std::isfinite(c->mu)) return fail(KF_ERR_INVALID, "mu must be positive");
    if (!(c->max_weight >= 1.0f) || c->max_weight > 32767.0f)
        return fail(KF_ERR_INVALID, "max_weight must be in [1, 32767]");
    if (!(c->near_plane >= 0.0f) || !(c->far_plane > c->near_plane) || c->far_plane > 1e4f)
        return fail(KF_ERR_INVALID, "need 0 <= near_plane < far_plane <= 1e4");
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(c->cam[k])) return fail(KF_ERR_INVALID, "camera parameters must be finite");
    if (!(c->cam[0] > 0.0f) || !(c->cam[1] > 0.0f)) return fail(KF_ERR_INVALID, "fx, fy must be positive");
    if (c->storage < 0 || c->storage > 2) return fail(KF_ERR_INVALID, "unknown storage");
    if (c->reduce_rule < 0 || c->reduce_rule > 1) return fail(KF_ERR_INVALID, "unknown reduce_rule");
    // stepSize() (gFusion.h:72-81) must advance t at every sample of the march
    const float minxyz = std::min(std::min(c->volume_dim[0], c->volume_dim[1]), c->volume_dim[2]);
    const float maxxyz = (float)std::max(std::max(c->volume_size[0], c->volume_size[1]), c->volume_size[2]);
    if (!(minxyz / maxxyz >= 1e-5f)) return fail(KF_ERR_INVALID, "voxel step below 1e-5 m");
    if (!(0.75f * c->mu >= 1e-5f)) return fail(KF_ERR_INVALID, "0.75*mu below 1e-5 m");
    // every t += step of the ray march must change t up to far_plane
    if (!(minxyz / maxxyz >= c->far_plane * 2e-6f) || !(0.75f * c->mu >= c->far_plane * 2e-6f))
        return fail(KF_ERR_INVALID, "ray step too small for far_plane");
    return KF_OK;
}

}  // namespace

namespace {

size_t session_bytes(const kf_config *c)
{
    const size_t p0 = (size_t)c->width * c->height;
    const size_t p1 = (size_t)(c->width >> 1) * (c->height >> 1);
    const size_t p2 = (size_t)(c->width >> 2) * (c->height >> 2);
    const size_t vox = (size_t)c->volume_size[0] * c->volume_size[1] * c->volume_size[2];
    size_t b = 0;
    b += p0 * 4 * 2 + p0 * 2 + p0 * 4;             // depth0, depth_m + staging
    b += 2 * 16 * (p0 + p1 + p2);                  // vertex + normal pyramids
    b += 2 * 16 * p0;                              // reference maps
    b += vox * 2 * voxel_plane_bytes(c->storage);  // TSDF value + weight planes
    b += sizeof(kf::DevState) + 2 * (size_t)(c->height + 8) * 32 * (4 + 8) + 64;  // state, ICP partial rows (plain, tagged)
    b += (size_t)((c->width + 7) / 8) * ((c->height + 7) / 8) * 8;  // depth tiles
    b += (size_t)kf::kPoseLibraryMax * 64 + 2 * 256 * 32 * 4;       // pose library, TrackSDF partials
    b += (size_t)(c->height + (c->height >> 1) + (c->height >> 2)) * 32;  // TrackSDF record slots
    const size_t nb = (size_t)((c->volume_size[0] + 7) / 8) * ((c->volume_size[1] + 7) / 8) *
                      ((c->volume_size[2] + 7) / 8);
    b += 4 * nb;                                   // brick flags, dilated, zero, coarse (bound)
    b += (size_t)c->volume_size[0] * c->volume_size[1] * 24 + 16;  // integrate's per-column set-up
    b += kf::ray_sched_bytes(((c->width + 15) / 16) * ((c->height + 15) / 16));  // hybrid raycast schedule
    return b;
}

void drop_graph(kf_session *s)
{
    if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->graph_exec = nullptr;
    s->graph = nullptr;
}

void free_all(kf_session *s)
{
    auto f = [](void *p) {
        if (p) (void)hipFree(p);
    };
    drop_graph(s);
    f(s->depth0); f(s->depth_m); f(s->stage_u16); f(s->stage_f32); f(s->bflags); f(s->dflags); f(s->cflags); f(s->zflags); f(s->dtiles); f(s->col_pos); f(s->col_range); f(s->col_dl); f(s->ray_sched);
    for (int l = 0; l < 3; ++l) {
        f(s->vert[l]);
        f(s->norm[l]);
    }
    f(s->refV); f(s->refN); f(s->tsdf); f(s->weight); f(s->st); f(s->block_partials); f(s->icp_rows); f(s->counts);
    f(s->dbg_result); f(s->dbg_error); f(s->dbg_J);
    f(s->library); f(s->sdf_partials); f(s->sdf_slots);
    f(s->mc_tab); f(s->mc_count); f(s->mc_offset); f(s->mc_part); f(s->mc_total); f(s->mc_verts); f(s->mip); f(s->mip_zflags);
    for (int i = 0; i < 8; ++i)
        if (s->tev[i]) (void)hipEventDestroy(s->tev[i]);
    for (int k = 0; k < 2; ++k) {
        if (s->h_stage_ev[k]) (void)hipEventSynchronize(s->h_stage_ev[k]);
        if (s->h_stage[k]) (void)hipHostFree(s->h_stage[k]);
        if (s->h_stage_ev[k]) (void)hipEventDestroy(s->h_stage_ev[k]);
    }
    if (s->own_stream && s->stream) (void)hipStreamDestroy(s->stream);
}

int level_w(const kf_session *s, int l) { return s->cfg.width >> l; }
int level_h(const kf_session *s, int l) { return s->cfg.height >> l; }

kf::PyramidArgs pyramid_args(kf_session *s)
{
    kf::PyramidArgs a{};
    a.w = s->cfg.width;
    a.h = s->cfg.height;
    for (int k = 0; k < 4; ++k) a.cam[k] = s->cfg.cam[k];
    a.tiles_x = (a.w + 7) / 8;
    a.tiles_y = (a.h + 7) / 8;
    // levels 2, 1, 0 in grid order: the coarse levels' threads average 16
    // and 4 input pixels per depth and take longest, so they start first
    int start = 0;
    for (int l = 2; l >= 0; --l) {
        a.block_start[l] = start;
        // level 0: one wavefront per 8x8 tile, four per block
        start += l == 0 ? (a.tiles_x * a.tiles_y + 3) / 4 : (level_w(s, l) * level_h(s, l) + 255) / 256;
        a.vert[l] = s->vert[l];
        a.norm[l] = s->norm[l];
    }
    a.blocks_total = start;
    a.depth0 = s->depth0;
    a.depth_m = s->depth_m;
    a.dtiles = s->dtiles;
    return a;
}

// target_segments: column segments per launch for this session alone; a
// batch, whose launch holds every member's, asks for fewer per session
kf::IntegrateArgs integrate_args(kf_session *s, int gated, long target_segments = 2097152)
{
    const kf_config &c = s->cfg;
    kf::IntegrateArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.N[k] = c.volume_size[k];
        a.dim[k] = c.volume_dim[k];
    }
    for (int k = 0; k < 4; ++k) a.cam[k] = c.cam[k];
    a.mu = c.mu;
    a.max_weight = c.max_weight;
    a.depth_m = s->depth_m;
    a.dw = c.width;
    a.dh = c.height;
    a.tsdf = s->tsdf;
    a.weight = s->weight;
    // enough z segments that the launch holds ~2M column segments (64-voxel
    // segments at 512^3: measured 106 us against 115 for 128 and 124 for 32)
    const long cols = (long)c.volume_size[0] * c.volume_size[1];
    int segs = (int)std::min<long>(std::max(1, c.volume_size[2] / 16),
                                   std::max<long>(1, (target_segments + cols - 1) / cols));
    int zlen = (c.volume_size[2] + segs - 1) / segs;
    if (s->int_zlen > 0) zlen = s->int_zlen;  // tuned voxels per thread
    zlen = (zlen + 7) / 8 * 8;
    a.zlen = std::max(zlen, 8);
    a.rows = s->int_rows;
    a.gated = gated;
    a.bflags = s->bflags;
    a.zflags = s->zflags;
    a.dtiles = s->dtiles;
    a.tiles_x = (c.width + 7) / 8;
    a.tiles_y = (c.height + 7) / 8;
    for (int k = 0; k < 3; ++k) a.NB[k] = s->NB[k];
    return a;
}

// lanes: rays per lane group (1 or 4); -1 takes the session's option
// the session's hybrid raycast schedule (tiles = 0 when the hybrid schedule
// is off: KF_OPT_RAYCAST_HYBRID 0 or four lanes per ray)
kf::RaySched ray_sched(kf_session *s)
{
    kf::RaySched r{};
    if (!s->ray_hybrid || s->ray_lanes != 1 || !s->ray_sched) return r;
    const int nt = ((s->cfg.width + 15) / 16) * ((s->cfg.height + 15) / 16);
    r.tile_cost = (uint32_t *)s->ray_sched;
    r.heavy = (int *)(s->ray_sched + 4 * (size_t)nt);
    r.heavy_flag = s->ray_sched + 4 * ((size_t)nt + kf::kRayHeavyMax + 1);
    r.tiles = nt;
    r.heavy_max = kf::kRayHeavyMax;
    return r;
}

kf::RaycastArgs raycast_args(kf_session *s, int lanes = -1, int xcd_group = -1)
{
    const kf_config &c = s->cfg;
    kf::RaycastArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.N[k] = c.volume_size[k];
        a.dim[k] = c.volume_dim[k];
    }
    for (int k = 0; k < 4; ++k) a.cam[k] = c.cam[k];
    a.near_plane = c.near_plane;
    a.far_plane = c.far_plane;
    // stepSize(): include/gFusion.h:72-81 ; largeStep = 0.75 * mu (gFusion.cpp:1307)
    const float minxyz = std::min(std::min(c.volume_dim[0], c.volume_dim[1]), c.volume_dim[2]);
    const float maxxyz = std::max(std::max((float)c.volume_size[0], (float)c.volume_size[1]), (float)c.volume_size[2]);
    a.step = minxyz / maxxyz;
    a.large_step = 0.75f * c.mu;
    a.tsdf = s->tsdf;
    a.w = c.width;
    a.h = c.height;
    a.lanes = lanes < 0 ? s->ray_lanes : lanes;
    const int tile = a.lanes == 4 ? 8 : 16;  // 256 threads: 256 or 64 rays
    a.tiles_x = (c.width + tile - 1) / tile;
    a.tiles_y = (c.height + tile - 1) / tile;
    a.xcd_group = std::max(1, xcd_group < 0 ? s->xcd_group : xcd_group);
    a.blocks = (a.tiles_x * a.tiles_y + 8 * a.xcd_group - 1) / (8 * a.xcd_group) * (8 * a.xcd_group);
    if (a.lanes == 1) {
        const kf::RaySched rs = ray_sched(s);
        if (rs.tiles > 0) {
            a.tile_cost = rs.tile_cost;
            a.heavy = rs.heavy;
            a.heavy_flag = rs.heavy_flag;
            a.heavy_max = rs.heavy_max;
        }
    }
    // 32-bit byte offsets, and N0 N1 < 2^24 for the 24-bit index products
    a.off32 = s->voxels * voxel_plane_bytes(c.storage) <= (size_t)0xFFFFFFFFu &&
                      (size_t)c.volume_size[0] * c.volume_size[1] < ((size_t)1 << 24)
                  ? 1
                  : 0;
    a.refV = s->refV;
    a.refN = s->refN;
    a.dflags = s->dflags;
    a.cflags = s->cflags;
    for (int k = 0; k < 3; ++k) {
        a.NB[k] = s->NB[k];
        a.NC[k] = s->NC[k];
    }
    a.skip = s->skip_empty;
    a.dim_pow2 = 1;
    for (int k = 0; k < 3; ++k) {
        int e;
        const float m = std::frexp(c.volume_dim[k], &e);
        // a power of two in 2^-32 .. 2^32 (e - 1 is its exponent)
        if (m != 0.5f || e - 1 < -32 || e - 1 > 32) a.dim_pow2 = 0;
        a.vscale[k] = (float)c.volume_size[k] * (1.0f / c.volume_dim[k]);  // exact when dim_pow2
    }
    return a;
}

kf::BrickArgs brick_args(kf_session *s)
{
    kf::BrickArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.N[k] = s->cfg.volume_size[k];
        a.NB[k] = s->NB[k];
        a.NC[k] = s->NC[k];
    }
    a.tsdf = s->tsdf;
    a.weight = s->weight;
    a.flags = s->bflags;
    a.dflags = s->dflags;
    a.cflags = s->cflags;
    a.zflags = s->zflags;
    return a;
}

// exact brick flags (occupancy and zero) from both planes, then the dilated
// copy; incremental: read only bricks whose zero flag is set (valid while
// every writer since the last rebuild has kept the flags as supersets: the
// frame loop's periodic rebuild, not after kf_write_tsdf)
int rebuild_bricks(kf_session *s, int incremental = 0)
{
    kf::BrickArgs a = brick_args(s);
    a.incremental = incremental;
    hipError_t e = kf::launch_brick_rebuild(a, s->cfg.storage, s->stream);
    if (e == hipSuccess) e = kf::launch_brick_dilate(a, s->stream);
    if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("brick rebuild: ") + hipGetErrorString(e));
    return KF_OK;
}

hipEvent_t tev(kf_session *s, int i) { return s->timing ? s->tev[i] : nullptr; }

int run_pyramid_from_stage(kf_session *s, const void *dev, int is_u16)
{
    kf::PyramidArgs a = pyramid_args(s);
    // one more block (the first) selects the hybrid raycast's next heavy
    // tiles from the session's earlier raycasts
    a.sched = ray_sched(s);
    if (a.sched.tiles > 0) {
        for (int l = 0; l < 3; ++l) a.block_start[l] += 1;
        a.blocks_total += 1;
    }
    hipError_t e = is_u16 ? kf::launch_pyramid_u16(a, (const uint16_t *)dev, s->stream, tev(s, 0), tev(s, 1))
                          : kf::launch_pyramid_f32(a, (const float *)dev, s->stream, tev(s, 0), tev(s, 1));
    if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("pyramid launch: ") + hipGetErrorString(e));
    return KF_OK;
}

kf::IcpArgs icp_args(kf_session *s, int level, int first, int save_pose)
{
    kf::IcpArgs a{};
    a.vert = s->vert[level];
    a.norm = s->norm[level];
    a.w = level_w(s, level);
    a.h = level_h(s, level);
    a.refV = s->refV;
    a.refN = s->refN;
    a.rw = s->cfg.width;
    a.rh = s->cfg.height;
    for (int k = 0; k < 4; ++k) a.cam[k] = s->cfg.cam[k];
    a.dist_th = s->cfg.dist_threshold;
    a.normal_th = s->cfg.normal_threshold;
    a.first = first;
    a.save_pose = save_pose;
    a.rule = s->cfg.reduce_rule;
    a.block_partials = s->block_partials;
    a.library = s->library;
    a.lib_max = kf::kPoseLibraryMax;
    // blocks per row group: about kIcpRowsPerBlock rows of the group per block
    const int rows_per_group = (a.h + 7) / 8;
    a.rows_step = std::max(1, (rows_per_group + s->icp_rows_per_block - 1) / s->icp_rows_per_block);
    // getReduction's last-row rule reads only row group 7 (gFusion.cpp:1143-1149):
    // the other groups' sums are not needed unless they are to be inspected
    a.group_only = (s->cfg.reduce_rule == KF_RULE_LAST_ROW && !s->icp_all_groups && !s->debug) ? 7 : -1;
    a.pose_fp64 = s->pose_fp64;
    // the reference-order mode keeps the library cos / sin (its bits are the
    // oracle's); the other modes, whose sums already differ in order, take
    // the series
    a.fast_trig = s->icp_ref_order ? 0 : 1;
    // small row groups: one block takes all rows of a group (the same
    // partial rows whether one group or all eight are summed)
    if (rows_per_group <= s->icp_solo_rows) a.rows_step = 1;
    return a;
}

int run_icp(kf_session *s, int level, int first, int save_pose, hipEvent_t t0 = nullptr, int finish = 0,
            int frame = 0, hipEvent_t t1 = nullptr)
{
    kf::IcpArgs a = icp_args(s, level, first, save_pose);
    a.finish = finish;
    a.frame = frame;
    if (s->debug || s->icp_ref_order) {
        a.dbg_result = s->dbg_result;
        a.dbg_error = s->dbg_error;
        a.dbg_J = s->dbg_J;
        s->dbg_level = level;
    }
    hipError_t e = s->icp_ref_order ? kf::launch_icp_ref_iteration(a, s->st, s->stream, t0, t1)
                                    : kf::launch_icp_iteration(a, s->st, s->stream, t0, t1);
    if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("icp launch: ") + hipGetErrorString(e));
    return KF_OK;
}

// Track() loop (gFusion.cpp:661-709): levels coarse -> fine, iterations[level]
// each; the early exit is a device flag, so every launch is queued up front.
// The last launch also applies the accept / reject (and the frame gate).
int run_track(kf_session *s, int frame)
{
    if (s->icp_persistent && !s->debug && !s->icp_ref_order) {
        // one launch for every level and iteration when its grid fits the chip
        kf::IcpTrackArgs a{};
        int nblk = 0, total = 0;
        for (int l = 0; l < 3; ++l) {
            a.lv[l] = icp_args(s, l, 0, 0);
            a.iters[l] = s->cfg.iterations[l];
            total += a.iters[l];
            if (a.iters[l] > 0) nblk = std::max(nblk, kf::icp_blocks(a.lv[l]));
        }
        if (total > 0 && nblk <= s->cus) {
            a.frame = frame;
            a.rows[0] = s->icp_rows;
            a.rows[1] = s->icp_rows + (size_t)(s->cfg.height + 8) * 32;
            a.spin_limit = 1u << 22;
            hipError_t e = kf::launch_icp_track(a, s->st, nblk, s->stream, tev(s, 2), tev(s, 3));
            if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("icp launch: ") + hipGetErrorString(e));
            return KF_OK;
        }
    }
    int last_level = -1;  // levels run 2 -> 0: the last one with iterations
    for (int level = 0; level <= 2 && last_level < 0; ++level)
        if (s->cfg.iterations[level] > 0) last_level = level;
    bool first_any = true;
    for (int level = 2; level >= 0; --level)
        for (int it = 0; it < s->cfg.iterations[level]; ++it) {
            const int fin = level == last_level && it == s->cfg.iterations[level] - 1;
            int rc = run_icp(s, level, it == 0, first_any ? 1 : 0, first_any ? tev(s, 2) : nullptr, fin, frame,
                             fin ? tev(s, 3) : nullptr);
            if (rc) return rc;
            first_any = false;
        }
    if (first_any) {
        // no iteration configured: AE stays 0 -> rejected, pose unchanged
        hipError_t e = kf::launch_state_set(s->st, nullptr, 0, s->stream);
        if (e == hipSuccess) e = kf::launch_icp_finish(s->st, frame, s->library, kf::kPoseLibraryMax, s->stream);
        if (e != hipSuccess) return fail(KF_ERR_HIP, hipGetErrorString(e));
    }
    return KF_OK;
}

// TrackSDF / recoverPose (gFusion.cpp:835-1038, 1616-1653): one persistent
// launch; dilate: it first does the brick-flag dilation of the integrate
// before it (k_brick_dilate's work), saving that launch
int run_sdf(kf_session *s, int mode, int frame, int dilate = 0)
{
    kf::SdfArgs a{};
    a.dilate = dilate;
    if (dilate) a.bricks = brick_args(s);
    int base = 0;
    for (int l = 0; l < 3; ++l) {
        a.vert[l] = s->vert[l];
        a.w[l] = level_w(s, l);
        a.h[l] = level_h(s, l);
        a.iters[l] = s->cfg.iterations[l];
        a.slot_base[l] = base;
        base += a.h[l];
    }
    a.tsdf = s->tsdf;
    for (int k = 0; k < 3; ++k) {
        a.N[k] = s->cfg.volume_size[k];
        a.dim[k] = s->cfg.volume_dim[k];
    }
    a.rule = s->cfg.reduce_rule;
    a.mode = mode;
    a.frame = frame;
    a.partials = s->sdf_partials;
    a.slots = s->sdf_slots;
    a.library = s->library;
    a.spin_limit = 1u << 22;
    hipError_t e = kf::launch_sdf_track(a, s->cfg.storage, s->st, s->sdf_blocks, s->stream);
    if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("TrackSDF launch: ") + hipGetErrorString(e));
    return KF_OK;
}

// The tracking step of the frame block (kinect.cpp:666-674): Track() and / or
// TrackSDF(); the last one run applies the frame gate.
int run_tracking(kf_session *s, int frame)
{
    int rc = KF_OK;
    if (s->track_mode & KF_TRACK_POINT_TO_PLANE) rc = run_track(s, frame && !(s->track_mode & KF_TRACK_POINT_TO_VOLUME));
    if (!rc && (s->track_mode & KF_TRACK_POINT_TO_VOLUME)) rc = run_sdf(s, kf::kSdfTrack, frame);
    return rc;
}

int run_integrate(kf_session *s, int gated, int dilate = 1)
{
    kf::IntegrateArgs a = integrate_args(s, gated);
    if (s->fast_math && s->int_columns) {
        a.col_pos = s->col_pos;
        a.col_range = s->col_range;
        a.col_dl = s->col_dl;
    }
    hipError_t e = kf::launch_integrate(a, s->cfg.storage, s->fast_math, s->st, s->stream, tev(s, 4), tev(s, 5));
    if (e == hipSuccess && dilate) e = kf::launch_brick_dilate(brick_args(s), s->stream);
    if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("integrate launch: ") + hipGetErrorString(e));
    return KF_OK;
}

int run_raycast(kf_session *s);

// Everything of the frame block after the pyramid: ICP with the frame gate,
// gated integrate (+ dilated brick flags), raycast.  Replayed from a hipGraph
// captured on first use (the launches and their arguments never change for a
// session); brick flags are rebuilt exactly every rebuild_period frames.
int run_frame_direct(kf_session *s)
{
    int rc = run_tracking(s, 1);
    // with recoverPose on, its launch does the integrate's flag dilation
    if (!rc) rc = run_integrate(s, 1, s->recover ? 0 : 1);
    // if (!tracked) recoverPose() (kinect.cpp:686-689): only the dilation
    // unless the tracking failed with a non-empty pose library
    if (!rc && s->recover) rc = run_sdf(s, kf::kSdfRecoverIfLost, 0, 1);
    if (!rc) rc = run_raycast(s);
    return rc;
}

int run_frame_body(kf_session *s)
{
    int rc = KF_OK;
    if (s->use_graph && !s->graph_failed && !s->debug && !s->timing) {
        if (!s->graph_exec) {
            hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                rc = run_frame_direct(s);
                hipGraph_t g = nullptr;
                hipError_t e2 = hipStreamEndCapture(s->stream, &g);
                if (rc == KF_OK && e2 == hipSuccess && g) {
                    s->graph = g;
                    if (hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
                        drop_graph(s);
                        s->graph_failed = 1;
                    }
                } else {
                    if (g) (void)hipGraphDestroy(g);
                    s->graph_failed = 1;
                }
            } else {
                s->graph_failed = 1;
            }
            (void)hipGetLastError();
        }
        if (s->graph_exec) {
            KF_HIP(hipGraphLaunch(s->graph_exec, s->stream));
        } else {
            rc = run_frame_direct(s);
        }
    } else {
        rc = run_frame_direct(s);
    }
    if (rc) return rc;
    s->frames++;
    if (s->rebuild_period > 0 && s->frames % s->rebuild_period == 0) rc = rebuild_bricks(s, 1);
    return rc;
}

int run_raycast(kf_session *s)
{
    kf::RaycastArgs a = raycast_args(s);
    const hipError_t e = kf::launch_raycast(a, s->cfg.storage, s->st, s->stream, tev(s, 6), tev(s, 7));
    if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("raycast launch: ") + hipGetErrorString(e));
    return KF_OK;
}

// per-pixel track output of one ICP iteration ({result, error, J[6]} of
// track.cs:18-27 as planes), for kf_set_debug and the reference-order Track()
int alloc_track_records(kf_session *s)
{
    if (s->dbg_result) return KF_OK;
    const size_t p0 = (size_t)s->cfg.width * s->cfg.height;
    if (hipMalloc((void **)&s->dbg_result, p0 * 4) != hipSuccess ||
        hipMalloc((void **)&s->dbg_error, p0 * 4) != hipSuccess || hipMalloc((void **)&s->dbg_J, p0 * 24) != hipSuccess)
        return fail(KF_ERR_OOM, "per-pixel track records");
    KF_HIP(hipMemsetAsync(s->dbg_result, 0, p0 * 4, s->stream));
    KF_HIP(hipMemsetAsync(s->dbg_error, 0, p0 * 4, s->stream));
    KF_HIP(hipMemsetAsync(s->dbg_J, 0, p0 * 24, s->stream));
    return KF_OK;
}

int map_ptr(kf_session *s, int which, int level, void **ptr, size_t *bytes)
{
    if (level < 0 || level > 2) return fail(KF_ERR_INVALID, "level must be 0..2");
    const size_t n = (size_t)level_w(s, level) * level_h(s, level);
    switch (which) {
    case KF_MAP_DEPTH:
        if (level != 0) return fail(KF_ERR_INVALID, "only depth level 0 is stored");
        *ptr = s->depth0; *bytes = n * 4; return KF_OK;
    case KF_MAP_VERTEX: *ptr = s->vert[level]; *bytes = n * 16; return KF_OK;
    case KF_MAP_NORMAL: *ptr = s->norm[level]; *bytes = n * 16; return KF_OK;
    case KF_MAP_REF_VERTEX:
        if (level != 0) return fail(KF_ERR_INVALID, "reference maps are level 0");
        *ptr = s->refV; *bytes = n * 16; return KF_OK;
    case KF_MAP_REF_NORMAL:
        if (level != 0) return fail(KF_ERR_INVALID, "reference maps are level 0");
        *ptr = s->refN; *bytes = n * 16; return KF_OK;
    case KF_MAP_INTENSITY:
        if (level != 0) return fail(KF_ERR_INVALID, "the intensity projection is level 0");
        if (!s->mip) return fail(KF_ERR_STATE, "kf_intensity_projection has not run");
        *ptr = s->mip; *bytes = n * 16; return KF_OK;
    }
    return fail(KF_ERR_INVALID, "unknown map");
}

// Host frame -> device staging buffer through a pinned slot: the caller's
// buffer is copied on the host (it may be reused as soon as the call
// returns, as depthToVertex's float* is), then sent by DMA on the session
// stream.  A slot is refilled only after its previous DMA completed.
int upload_host(kf_session *s, const void *src, size_t bytes, void *dst)
{
    const size_t cap = (size_t)s->cfg.width * s->cfg.height * 4;
    if (bytes > cap) return fail(KF_ERR_INVALID, "upload larger than a depth frame");
    if (!s->h_stage[0]) {
        for (int k = 0; k < 2; ++k) {
            if (hipHostMalloc(&s->h_stage[k], cap, hipHostMallocDefault) != hipSuccess)
                return fail(KF_ERR_OOM, "pinned staging buffer");
            if (hipEventCreateWithFlags(&s->h_stage_ev[k], hipEventDisableTiming) != hipSuccess)
                return fail(KF_ERR_HIP, "hipEventCreate failed");
        }
    }
    const int k = s->h_stage_next;
    if (s->h_stage_pending[k]) KF_HIP(hipEventSynchronize(s->h_stage_ev[k]));
    s->h_stage_pending[k] = false;
    std::memcpy(s->h_stage[k], src, bytes);
    KF_HIP(hipMemcpyAsync(dst, s->h_stage[k], bytes, hipMemcpyHostToDevice, s->stream));
    KF_HIP(hipEventRecord(s->h_stage_ev[k], s->stream));
    s->h_stage_pending[k] = true;
    s->h_stage_next = k ^ 1;
    return KF_OK;
}

}  // namespace

int kf_fail(int code, const char *msg) { return fail(code, msg); }

extern "C" {

const char *kf_last_error(void) { return g_err.c_str(); }
const char *kf_version(void) { return "kfusion-gfx950 0.1"; }

int kf_config_default(kf_config *c)
{
    if (!c) return fail(KF_ERR_INVALID, "config is NULL");
    std::memset(c, 0, sizeof(*c));
    for (int k = 0; k < 3; ++k) {
        c->volume_size[k] = 256;   // gFusion.h:57
        c->volume_dim[k] = 1.0f;   // gFusion.h:58
        c->iterations[k] = 5;      // gFusion.h:66-68
    }
    c->mu = 0.05f;
    c->max_weight = 100.0f;
    c->near_plane = 0.1f;
    c->far_plane = 3.0f;
    c->dist_threshold = 0.05f;
    c->normal_threshold = 0.8f;
    c->cam[0] = 365.5f;
    c->cam[1] = 365.5f;
    c->cam[2] = 255.5f;
    c->cam[3] = 211.5f;
    c->width = 512;
    c->height = 424;
    c->storage = KF_STORAGE_RG16I;
    c->reduce_rule = KF_RULE_LAST_ROW;
    return KF_OK;
}

int kf_session_bytes(const kf_config *cfg, size_t *bytes)
{
    int rc = validate(cfg);
    if (rc) return rc;
    if (!bytes) return fail(KF_ERR_INVALID, "bytes is NULL");
    *bytes = session_bytes(cfg);
    return KF_OK;
}

// Launch shapes picked by the autotuner (oglfusionkinectv2_amd/autotune.py)
// per storage type and volume size; the session takes the row of its storage
// with the nearest size (largest side), RG16I the FP16 rows (both 4 B per voxel).
namespace {
struct TunedRow {
    int storage, volume, int_zlen, int_rows, xcd_group, ray_lanes, rebuild_period, int_columns, ray_hybrid;
};
static const TunedRow kTuned[] = {
#include "kf_tuned.inc"
    {-1, 0, 0, 4, 4, 1, 16, 1, 1},  // end marker
};

void apply_tuned(kf_session *s)
{
    const int st = s->cfg.storage == KF_STORAGE_F32 ? KF_STORAGE_F32 : KF_STORAGE_F16;
    const int n = std::max(std::max(s->cfg.volume_size[0], s->cfg.volume_size[1]), s->cfg.volume_size[2]);
    const TunedRow *best = nullptr;
    for (const TunedRow &r : kTuned) {
        if (r.storage != st) continue;
        if (!best || std::abs(r.volume - n) < std::abs(best->volume - n)) best = &r;
    }
    if (!best) return;
    s->int_zlen = best->int_zlen;
    s->int_rows = best->int_rows;
    s->xcd_group = best->xcd_group;
    s->ray_lanes = best->ray_lanes;
    s->rebuild_period = best->rebuild_period;
    s->int_columns = best->int_columns;
    s->ray_hybrid = best->ray_hybrid;
}
}  // namespace

int kf_session_create(const kf_config *cfg, void *stream, kf_session **out)
{
    int rc = validate(cfg);
    if (rc) return rc;
    if (!out) return fail(KF_ERR_INVALID, "out is NULL");
    *out = nullptr;
    kf_session *s = new (std::nothrow) kf_session();
    if (!s) return fail(KF_ERR_OOM, "host allocation failed");
    s->cfg = *cfg;
    apply_tuned(s);
    auto cleanup = [&](int code, const std::string &msg) {
        free_all(s);
        delete s;
        return fail(code, msg);
    };
    if (hipGetDevice(&s->device) != hipSuccess) return cleanup(KF_ERR_HIP, "no HIP device");
    if (kf::mesh_noalloc_setup(s->device) != KF_OK) return cleanup(KF_ERR_HIP, kf_last_error());
    if (stream) {
        s->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(KF_ERR_HIP, "hipStreamCreate failed");
        s->own_stream = true;
    }
    const size_t p0 = (size_t)cfg->width * cfg->height;
    s->voxels = (size_t)cfg->volume_size[0] * cfg->volume_size[1] * cfg->volume_size[2];
    const size_t vb = s->voxels * voxel_plane_bytes(cfg->storage);
    auto alloc = [&](void **p, size_t n) { return hipMalloc(p, n) == hipSuccess; };
    for (int k = 0; k < 3; ++k) {
        s->NB[k] = (cfg->volume_size[k] + 7) / 8;
        s->NC[k] = (s->NB[k] + 3) / 4;
    }
    s->nbricks = (size_t)s->NB[0] * s->NB[1] * s->NB[2];
    s->ncoarse = (size_t)s->NC[0] * s->NC[1] * s->NC[2];
    bool ok = alloc((void **)&s->depth0, p0 * 4) && alloc((void **)&s->depth_m, p0 * 4) &&
              alloc((void **)&s->bflags, s->nbricks) && alloc((void **)&s->dflags, s->nbricks) &&
              alloc((void **)&s->zflags, s->nbricks) &&
              alloc((void **)&s->dtiles, (size_t)((cfg->width + 7) / 8) * ((cfg->height + 7) / 8) * 8) &&
              alloc((void **)&s->cflags, s->ncoarse) &&
              alloc((void **)&s->col_pos, (size_t)cfg->volume_size[0] * cfg->volume_size[1] * 16) &&
              alloc((void **)&s->col_range, (size_t)cfg->volume_size[0] * cfg->volume_size[1] * 8) &&
              alloc((void **)&s->col_dl, 16) &&
              alloc((void **)&s->ray_sched, kf::ray_sched_bytes(((cfg->width + 15) / 16) * ((cfg->height + 15) / 16))) &&
              alloc((void **)&s->stage_u16, p0 * 2) &&
              alloc((void **)&s->stage_f32, p0 * 4) && alloc((void **)&s->refV, p0 * 16) &&
              alloc((void **)&s->refN, p0 * 16) && alloc(&s->tsdf, vb) && alloc(&s->weight, vb) &&
              alloc((void **)&s->st, sizeof(kf::DevState)) &&
              alloc((void **)&s->block_partials, 2 * (size_t)(cfg->height + 8) * 32 * 4) &&
              alloc((void **)&s->icp_rows, 2 * (size_t)(cfg->height + 8) * 32 * 8) &&
              alloc((void **)&s->counts, 8 * sizeof(unsigned long long)) &&
              alloc((void **)&s->library, (size_t)kf::kPoseLibraryMax * 64) &&
              alloc((void **)&s->sdf_partials, 2 * 256 * 32 * 4) &&
              alloc((void **)&s->sdf_slots, (size_t)(cfg->height + (cfg->height >> 1) + (cfg->height >> 2)) * 32);
    for (int l = 0; l < 3 && ok; ++l) {
        const size_t n = (size_t)(cfg->width >> l) * (cfg->height >> l);
        ok = alloc((void **)&s->vert[l], n * 16) && alloc((void **)&s->norm[l], n * 16);
    }
    if (!ok) return cleanup(KF_ERR_OOM, "hipMalloc failed (session needs " + std::to_string(session_bytes(cfg)) + " B)");
    // initial contents: zero maps / volume / state, identity pose
    bool z = hipMemsetAsync(s->tsdf, 0, vb, s->stream) == hipSuccess &&
             hipMemsetAsync(s->weight, 0, vb, s->stream) == hipSuccess &&
             hipMemsetAsync(s->refV, 0, p0 * 16, s->stream) == hipSuccess &&
             hipMemsetAsync(s->refN, 0, p0 * 16, s->stream) == hipSuccess &&
             hipMemsetAsync(s->depth0, 0, p0 * 4, s->stream) == hipSuccess &&
             hipMemsetAsync(s->depth_m, 0, p0 * 4, s->stream) == hipSuccess &&
             hipMemsetAsync(s->bflags, 0, s->nbricks, s->stream) == hipSuccess &&
             hipMemsetAsync(s->dflags, 0, s->nbricks, s->stream) == hipSuccess &&
             hipMemsetAsync(s->zflags, 0, s->nbricks, s->stream) == hipSuccess &&
             hipMemsetAsync(s->dtiles, 0, (size_t)((cfg->width + 7) / 8) * ((cfg->height + 7) / 8) * 8,
                            s->stream) == hipSuccess &&
             hipMemsetAsync(s->cflags, 0, s->ncoarse, s->stream) == hipSuccess &&
             hipMemsetAsync(s->ray_sched, 0, kf::ray_sched_bytes(((cfg->width + 15) / 16) * ((cfg->height + 15) / 16)),
                            s->stream) == hipSuccess &&
             hipMemsetAsync(s->st, 0, sizeof(kf::DevState), s->stream) == hipSuccess &&
             hipMemsetAsync(s->icp_rows, 0, 2 * (size_t)(cfg->height + 8) * 32 * 8, s->stream) == hipSuccess &&
             hipMemsetAsync(s->library, 0, (size_t)kf::kPoseLibraryMax * 64, s->stream) == hipSuccess &&
             hipMemsetAsync(s->sdf_slots, 0, (size_t)(cfg->height + (cfg->height >> 1) + (cfg->height >> 2)) * 32,
                            s->stream) == hipSuccess;
    for (int l = 0; l < 3 && z; ++l) {
        const size_t n = (size_t)(cfg->width >> l) * (cfg->height >> l);
        z = hipMemsetAsync(s->vert[l], 0, n * 16, s->stream) == hipSuccess &&
            hipMemsetAsync(s->norm[l], 0, n * 16, s->stream) == hipSuccess;
    }
    {
        // persistent TrackSDF grid: every block must be resident at once
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess || cus < 8)
            return cleanup(KF_ERR_HIP, "cannot query the compute unit count");
        s->sdf_blocks = (cus < 256 ? cus : 256) & ~7;
        s->cus = cus;
    }
    const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (!z || kf::launch_state_set(s->st, I, kf::kStateResetPose, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return cleanup(KF_ERR_HIP, "session initialisation failed");
    *out = s;
    return KF_OK;
}

int kf_session_destroy(kf_session *s)
{
    if (!s) return KF_OK;
    // batches may still hold queued launches that read this session on their own streams
    (void)hipDeviceSynchronize();
    free_all(s);
    delete s;
    return KF_OK;
}

int kf_get_config(const kf_session *s, kf_config *cfg)
{
    if (!s || !cfg) return fail(KF_ERR_INVALID, "NULL argument");
    *cfg = s->cfg;
    return KF_OK;
}

int kf_reset_volume(kf_session *s)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    const size_t vb = s->voxels * voxel_plane_bytes(s->cfg.storage);
    KF_HIP(hipMemsetAsync(s->tsdf, 0, vb, s->stream));
    KF_HIP(hipMemsetAsync(s->weight, 0, vb, s->stream));
    KF_HIP(hipMemsetAsync(s->bflags, 0, s->nbricks, s->stream));
    KF_HIP(hipMemsetAsync(s->dflags, 0, s->nbricks, s->stream));
    KF_HIP(hipMemsetAsync(s->zflags, 0, s->nbricks, s->stream));
    KF_HIP(hipMemsetAsync(s->cflags, 0, s->ncoarse, s->stream));
    // the raycast's heavy-tile schedule starts over with the volume
    KF_HIP(hipMemsetAsync(s->ray_sched, 0,
                          kf::ray_sched_bytes(((s->cfg.width + 15) / 16) * ((s->cfg.height + 15) / 16)), s->stream));
    return KF_OK;
}

int kf_set_pose(kf_session *s, const float pose[16])
{
    if (!s || !pose) return fail(KF_ERR_INVALID, "NULL argument");
    KF_HIP(kf::launch_state_set(s->st, pose, 0, s->stream));
    return KF_OK;
}

int kf_reset(kf_session *s, const float pose[16])
{
    if (!s || !pose) return fail(KF_ERR_INVALID, "NULL argument");
    int rc = kf_reset_volume(s);
    if (rc) return rc;
    KF_HIP(kf::launch_state_set(s->st, pose, kf::kStateArmWindow | kf::kStateResetPose, s->stream));
    return run_raycast(s);
}

int kf_get_pose(kf_session *s, float pose[16])
{
    if (!s || !pose) return fail(KF_ERR_INVALID, "NULL argument");
    KF_HIP(hipMemcpyAsync(pose, s->st->pose, 64, hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_depth_to_vertex_f32(kf_session *s, const float *depth_mm)
{
    if (!s || !depth_mm) return fail(KF_ERR_INVALID, "NULL argument");
    const size_t n = (size_t)s->cfg.width * s->cfg.height;
    const int rc = upload_host(s, depth_mm, n * 4, s->stage_f32);
    return rc ? rc : run_pyramid_from_stage(s, s->stage_f32, 0);
}

int kf_depth_to_vertex_u16(kf_session *s, const uint16_t *depth_mm)
{
    if (!s || !depth_mm) return fail(KF_ERR_INVALID, "NULL argument");
    const size_t n = (size_t)s->cfg.width * s->cfg.height;
    const int rc = upload_host(s, depth_mm, n * 2, s->stage_u16);
    return rc ? rc : run_pyramid_from_stage(s, s->stage_u16, 1);
}

int kf_depth_to_vertex_u16_dev(kf_session *s, const uint16_t *depth_mm_dev)
{
    if (!s || !depth_mm_dev) return fail(KF_ERR_INVALID, "NULL argument");
    return run_pyramid_from_stage(s, depth_mm_dev, 1);
}

int kf_depth_to_vertex_f32_dev(kf_session *s, const float *depth_mm_dev)
{
    if (!s || !depth_mm_dev) return fail(KF_ERR_INVALID, "NULL argument");
    return run_pyramid_from_stage(s, depth_mm_dev, 0);
}

int kf_icp_iteration(kf_session *s, int level, int first, int save_pose)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    if (level < 0 || level > 2) return fail(KF_ERR_INVALID, "level must be 0..2");
    return run_icp(s, level, first ? 1 : 0, save_pose ? 1 : 0);
}

int kf_icp_finish(kf_session *s)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    KF_HIP(kf::launch_icp_finish(s->st, 0, s->library, kf::kPoseLibraryMax, s->stream));
    return KF_OK;
}

int kf_track(kf_session *s, int *tracked)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    int rc = run_track(s, 0);
    if (rc) return rc;
    if (tracked) {
        KF_HIP(hipMemcpyAsync(tracked, &s->st->tracked, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        KF_HIP(hipStreamSynchronize(s->stream));
    }
    return KF_OK;
}

int kf_track_frame(kf_session *s)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    return run_tracking(s, 1);
}

int kf_track_sdf(kf_session *s, int *tracked)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    int rc = run_sdf(s, kf::kSdfTrack, 0);
    if (rc || !tracked) return rc;
    int v[2];
    KF_HIP(hipMemcpyAsync(v, &s->st->sdf_tracked, 2 * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    *tracked = v[0];
    return KF_OK;
}

int kf_recover_pose(kf_session *s, int *found)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    int rc = run_sdf(s, kf::kSdfRecover, 0);
    if (rc || !found) return rc;
    KF_HIP(hipMemcpyAsync(found, &s->st->sdf_tracked, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_read_pose_library(kf_session *s, float *poses, int capacity, int *count)
{
    if (!s || !count) return fail(KF_ERR_INVALID, "NULL argument");
    int n = 0;
    KF_HIP(hipMemcpyAsync(&n, &s->st->lib_n, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    *count = n;
    if (poses && capacity > 0) {
        const int m = n < capacity ? n : capacity;
        if (m > 0) {
            KF_HIP(hipMemcpyAsync(poses, s->library, (size_t)m * 64, hipMemcpyDeviceToHost, s->stream));
            KF_HIP(hipStreamSynchronize(s->stream));
        }
    }
    return KF_OK;
}

int kf_read_sdf_state(kf_session *s, float pose_eig[16], float cum_twist[6], float partials[256], float *ae,
                      int *iterations_run, double *log)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    kf::DevState h;
    KF_HIP(hipMemcpyAsync(&h, s->st, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    if (h.sdf_error) return fail(KF_ERR_STATE, "a TrackSDF grid barrier timed out");
    if (pose_eig) std::memcpy(pose_eig, h.pose_eig, 64);
    if (cum_twist)
        for (int k = 0; k < 6; ++k) cum_twist[k] = (float)h.cum_twist[k];
    if (partials) std::memcpy(partials, h.sdf_partials, 256 * 4);
    if (ae) *ae = h.sdf_ae;
    if (iterations_run) *iterations_run = h.sdf_iters;
    if (log) std::memcpy(log, h.sdf_log, sizeof(h.sdf_log));
    return KF_OK;
}

int kf_write_pose_eig(kf_session *s, const float pose_eig[16])
{
    if (!s || !pose_eig) return fail(KF_ERR_INVALID, "NULL argument");
    KF_HIP(hipMemcpyAsync(s->st->pose_eig, pose_eig, 64, hipMemcpyHostToDevice, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_get_cum_twist(kf_session *s, float twist[6])
{
    if (!twist) return fail(KF_ERR_INVALID, "NULL argument");
    return kf_read_sdf_state(s, nullptr, twist, nullptr, nullptr, nullptr, nullptr);
}

int kf_intensity_projection(kf_session *s)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    if (!s->mip) {
        const size_t p0 = (size_t)s->cfg.width * s->cfg.height;
        if (hipMalloc((void **)&s->mip, p0 * 16) != hipSuccess) return fail(KF_ERR_OOM, "intensity image");
    }
    if (!s->mip_zflags && hipMalloc((void **)&s->mip_zflags, s->nbricks + s->ncoarse) != hipSuccess)
        return fail(KF_ERR_OOM, "intensity projection flags");
    kf::RaycastArgs a = raycast_args(s);
    hipError_t e = hipSuccess;
    if (a.skip) {
        // the zero flags dilated as the raycast's flags are: a brick's flag is
        // clear when it and its +1 neighbours hold only (0, 0)
        kf::BrickArgs z = brick_args(s);
        z.flags = s->zflags;
        z.dflags = s->mip_zflags;
        z.cflags = s->mip_zflags + s->nbricks;
        e = kf::launch_brick_dilate(z, s->stream);
        a.dflags = z.dflags;
        a.cflags = z.cflags;
    }
    if (e == hipSuccess) e = kf::launch_intensity_projection(a, s->cfg.storage, s->st, s->mip, s->stream);
    if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("intensity projection: ") + hipGetErrorString(e));
    return KF_OK;
}

int kf_integrate(kf_session *s, int mode)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    if (mode != KF_INTEGRATE_ALWAYS && mode != KF_INTEGRATE_GATED) return fail(KF_ERR_INVALID, "unknown mode");
    return run_integrate(s, mode == KF_INTEGRATE_GATED);
}

int kf_raycast(kf_session *s)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    return run_raycast(s);
}

int kf_frame_u16_dev(kf_session *s, const uint16_t *depth_mm_dev)
{
    if (!s || !depth_mm_dev) return fail(KF_ERR_INVALID, "NULL argument");
    int rc = run_pyramid_from_stage(s, depth_mm_dev, 1);
    return rc ? rc : run_frame_body(s);
}

int kf_frame_f32(kf_session *s, const float *depth_mm)
{
    int rc = kf_depth_to_vertex_f32(s, depth_mm);
    return rc ? rc : run_frame_body(s);
}

int kf_set_option(kf_session *s, int option, int value)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    switch (option) {
    case KF_OPT_FAST_MATH: s->fast_math = value ? 1 : 0; break;
    case KF_OPT_SKIP_EMPTY: s->skip_empty = value ? 1 : 0; break;
    case KF_OPT_GRAPH: s->use_graph = value ? 1 : 0; break;
    case KF_OPT_STAGE_TIMING:
        if (value && !s->tev[0])
            for (int i = 0; i < 8; ++i)
                if (hipEventCreate(&s->tev[i]) != hipSuccess) return fail(KF_ERR_HIP, "hipEventCreate failed");
        s->timing = value ? 1 : 0;
        break;
    case KF_OPT_BRICK_REBUILD_PERIOD:
        if (value < 0) return fail(KF_ERR_INVALID, "rebuild period must be >= 0");
        s->rebuild_period = value;
        break;
    case KF_OPT_RAYCAST_XCD_GROUP:
        if (value < 1 || value > 4096) return fail(KF_ERR_INVALID, "xcd group must be in [1, 4096]");
        s->xcd_group = value;
        break;
    case KF_OPT_ICP_ROWS_PER_BLOCK:
        if (value < 1 || value > 1024) return fail(KF_ERR_INVALID, "rows per block must be in [1, 1024]");
        s->icp_rows_per_block = value;
        break;
    case KF_OPT_ICP_ALL_GROUPS: s->icp_all_groups = value ? 1 : 0; break;
    case KF_OPT_RAYCAST_LANES:
        if (value != 1 && value != 4) return fail(KF_ERR_INVALID, "raycast lanes per ray must be 1 or 4");
        s->ray_lanes = value;
        break;
    case KF_OPT_TRACK_MODE:
        if (value < 1 || value > 3) return fail(KF_ERR_INVALID, "track mode must be 1, 2 or 3");
        s->track_mode = value;
        break;
    case KF_OPT_RECOVER: s->recover = value ? 1 : 0; break;
    case KF_OPT_ICP_PERSISTENT: s->icp_persistent = value ? 1 : 0; break;
    case KF_OPT_ICP_SOLO_ROWS:
        if (value < 0) return fail(KF_ERR_INVALID, "solo rows must be >= 0");
        s->icp_solo_rows = value;
        break;
    case KF_OPT_INTEGRATE_ZLEN:
        if (value < 0 || value > 4096 || value % 8) return fail(KF_ERR_INVALID, "integrate z length must be 0 or a multiple of 8 up to 4096");
        s->int_zlen = value;
        break;
    case KF_OPT_INTEGRATE_ROWS:
        if (value != 1 && value != 2 && value != 4) return fail(KF_ERR_INVALID, "integrate block rows must be 1, 2 or 4");
        s->int_rows = value;
        break;
    case KF_OPT_POSE_FP64: s->pose_fp64 = value ? 1 : 0; break;
    case KF_OPT_INTEGRATE_COLUMNS: s->int_columns = value ? 1 : 0; break;
    case KF_OPT_RAYCAST_HYBRID: s->ray_hybrid = value ? 1 : 0; break;
    case KF_OPT_RAYCAST_HEAVY_TILES: return fail(KF_ERR_INVALID, "the heavy-tile count is read-only");
    case KF_OPT_ICP_REFERENCE_ORDER:
        if (value) {
            const int rc = alloc_track_records(s);
            if (rc) return rc;
        }
        s->icp_ref_order = value ? 1 : 0;
        break;
    default: return fail(KF_ERR_INVALID, "unknown option");
    }
    drop_graph(s);  // captured launches carry the old arguments
    if (option == KF_OPT_ICP_ROWS_PER_BLOCK || option == KF_OPT_ICP_ALL_GROUPS || option == KF_OPT_ICP_SOLO_ROWS) {
        // the persistent Track() may now read partial rows that no launch of
        // the old layout wrote: clear their tags (a row is valid only when it
        // carries the current launch's tag, and a row left from an old layout
        // could carry a wrapped-around one)
        KF_HIP(hipMemsetAsync(s->icp_rows, 0, 2 * (size_t)(s->cfg.height + 8) * 32 * 8, s->stream));
    }
    return KF_OK;
}

int kf_get_option(kf_session *s, int option, int *value)
{
    if (!s || !value) return fail(KF_ERR_INVALID, "NULL argument");
    switch (option) {
    case KF_OPT_FAST_MATH: *value = s->fast_math; break;
    case KF_OPT_SKIP_EMPTY: *value = s->skip_empty; break;
    case KF_OPT_GRAPH: *value = s->use_graph && !s->graph_failed; break;
    case KF_OPT_STAGE_TIMING: *value = s->timing; break;
    case KF_OPT_BRICK_REBUILD_PERIOD: *value = s->rebuild_period; break;
    case KF_OPT_RAYCAST_XCD_GROUP: *value = s->xcd_group; break;
    case KF_OPT_ICP_ROWS_PER_BLOCK: *value = s->icp_rows_per_block; break;
    case KF_OPT_TRACK_MODE: *value = s->track_mode; break;
    case KF_OPT_ICP_ALL_GROUPS: *value = s->icp_all_groups; break;
    case KF_OPT_RAYCAST_LANES: *value = s->ray_lanes; break;
    case KF_OPT_RECOVER: *value = s->recover; break;
    case KF_OPT_ICP_PERSISTENT: *value = s->icp_persistent; break;
    case KF_OPT_ICP_SOLO_ROWS: *value = s->icp_solo_rows; break;
    case KF_OPT_ICP_REFERENCE_ORDER: *value = s->icp_ref_order; break;
    case KF_OPT_INTEGRATE_ZLEN: *value = s->int_zlen; break;
    case KF_OPT_INTEGRATE_ROWS: *value = s->int_rows; break;
    case KF_OPT_POSE_FP64: *value = s->pose_fp64; break;
    case KF_OPT_INTEGRATE_COLUMNS: *value = s->int_columns; break;
    case KF_OPT_RAYCAST_HYBRID: *value = s->ray_hybrid; break;
    case KF_OPT_RAYCAST_HEAVY_TILES: {
        const int nt = ((s->cfg.width + 15) / 16) * ((s->cfg.height + 15) / 16);
        int h = 0;
        KF_HIP(hipMemcpyAsync(&h, s->ray_sched + 4 * (size_t)nt, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        KF_HIP(hipStreamSynchronize(s->stream));
        *value = h;
        break;
    }
    default: return fail(KF_ERR_INVALID, "unknown option");
    }
    return KF_OK;
}

int kf_read_stage_times(kf_session *s, float ms[4])
{
    if (!s || !ms) return fail(KF_ERR_INVALID, "NULL argument");
    if (!s->timing) return fail(KF_ERR_STATE, "stage timing is off (KF_OPT_STAGE_TIMING)");
    KF_HIP(hipStreamSynchronize(s->stream));
    for (int k = 0; k < 4; ++k) {
        const hipError_t e = hipEventElapsedTime(&ms[k], s->tev[2 * k], s->tev[2 * k + 1]);
        if (e != hipSuccess) return fail(KF_ERR_STATE, "stage " + std::to_string(k) + " has not run with timing on");
    }
    return KF_OK;
}

int kf_rebuild_bricks(kf_session *s)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    return rebuild_bricks(s);
}

int kf_read_brick_flags(kf_session *s, uint8_t *occupancy_host, uint8_t *zero_host)
{
    if (!s || !occupancy_host || !zero_host) return fail(KF_ERR_INVALID, "NULL argument");
    KF_HIP(hipMemcpyAsync(occupancy_host, s->bflags, s->nbricks, hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipMemcpyAsync(zero_host, s->zflags, s->nbricks, hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_read_map(kf_session *s, int which, int level, void *host, size_t bytes)
{
    if (!s || !host) return fail(KF_ERR_INVALID, "NULL argument");
    void *p;
    size_t n;
    int rc = map_ptr(s, which, level, &p, &n);
    if (rc) return rc;
    if (bytes != n) return fail(KF_ERR_INVALID, "map size is " + std::to_string(n) + " bytes");
    KF_HIP(hipMemcpyAsync(host, p, n, hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_write_map(kf_session *s, int which, int level, const void *host, size_t bytes)
{
    if (!s || !host) return fail(KF_ERR_INVALID, "NULL argument");
    void *p;
    size_t n;
    int rc = map_ptr(s, which, level, &p, &n);
    if (rc) return rc;
    if (bytes != n) return fail(KF_ERR_INVALID, "map size is " + std::to_string(n) + " bytes");
    KF_HIP(hipMemcpyAsync(p, host, n, hipMemcpyHostToDevice, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_read_tsdf(kf_session *s, void *value, void *weight)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    const size_t vb = s->voxels * voxel_plane_bytes(s->cfg.storage);
    if (value) KF_HIP(hipMemcpyAsync(value, s->tsdf, vb, hipMemcpyDeviceToHost, s->stream));
    if (weight) KF_HIP(hipMemcpyAsync(weight, s->weight, vb, hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_write_tsdf(kf_session *s, const void *value, const void *weight)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    const size_t vb = s->voxels * voxel_plane_bytes(s->cfg.storage);
    if (value) KF_HIP(hipMemcpyAsync(s->tsdf, value, vb, hipMemcpyHostToDevice, s->stream));
    if (weight) KF_HIP(hipMemcpyAsync(s->weight, weight, vb, hipMemcpyHostToDevice, s->stream));
    int rc = rebuild_bricks(s);
    if (rc) return rc;
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_read_reduction(kf_session *s, float partials[256], float *ae)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    if (partials) KF_HIP(hipMemcpyAsync(partials, s->st->partials, 256 * 4, hipMemcpyDeviceToHost, s->stream));
    if (ae) KF_HIP(hipMemcpyAsync(ae, &s->st->ae, 4, hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_read_icp_solution(kf_session *s, double x[6])
{
    if (!s || !x) return fail(KF_ERR_INVALID, "NULL argument");
    KF_HIP(hipMemcpyAsync(x, s->st->x, 6 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_read_status(kf_session *s, int *tracked, int *iterations_run, int *counter, int *reset)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    kf::DevState h;
    KF_HIP(hipMemcpyAsync(&h, s->st, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    if (h.icp_error) return fail(KF_ERR_STATE, "a Track() grid barrier timed out (cleared by kf_reset)");
    if (tracked) *tracked = h.tracked;
    if (iterations_run) *iterations_run = h.iters_run;
    if (counter) *counter = h.counter;
    if (reset) *reset = h.reset;
    return KF_OK;
}

int kf_set_debug(kf_session *s, int record)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    if (record) {
        const int rc = alloc_track_records(s);
        if (rc) return rc;
    }
    s->debug = record ? 1 : 0;
    drop_graph(s);
    return KF_OK;
}

int kf_read_track_pixels(kf_session *s, int level, int32_t *result, float *error, float *J)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    if (!s->dbg_result) return fail(KF_ERR_STATE, "per-pixel recording is off (kf_set_debug)");
    if (level != s->dbg_level) return fail(KF_ERR_STATE, "last recorded iteration was at another level");
    const size_t n = (size_t)level_w(s, level) * level_h(s, level);
    if (result) KF_HIP(hipMemcpyAsync(result, s->dbg_result, n * 4, hipMemcpyDeviceToHost, s->stream));
    if (error) KF_HIP(hipMemcpyAsync(error, s->dbg_error, n * 4, hipMemcpyDeviceToHost, s->stream));
    if (J) KF_HIP(hipMemcpyAsync(J, s->dbg_J, n * 24, hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_synchronize(kf_session *s)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    KF_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

int kf_get_stream(kf_session *s, void **stream)
{
    if (!s || !stream) return fail(KF_ERR_INVALID, "NULL argument");
    *stream = (void *)s->stream;
    return KF_OK;
}

int kf_raycast_stats(kf_session *s, uint64_t stats[6])
{
    if (!s || !stats) return fail(KF_ERR_INVALID, "NULL argument");
    KF_HIP(hipMemsetAsync(s->counts, 0, 6 * sizeof(unsigned long long), s->stream));
    kf::RaycastArgs a = raycast_args(s);
    a.stats = s->counts;
    KF_HIP(kf::launch_raycast(a, s->cfg.storage, s->st, s->stream));
    unsigned long long h[6];
    KF_HIP(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    for (int k = 0; k < 6; ++k) stats[k] = h[k];
    return KF_OK;
}

int kf_raycast_ray_counts(kf_session *s, uint32_t *counts_host)
{
    if (!s || !counts_host) return fail(KF_ERR_INVALID, "NULL argument");
    const size_t np = (size_t)s->cfg.width * s->cfg.height;
    uint32_t *d = nullptr;
    KF_HIP(hipMalloc(&d, np * sizeof(uint32_t)));
    kf::RaycastArgs a = raycast_args(s, 1);
    a.stats = s->counts;
    a.ray_counts = d;
    hipError_t e = hipMemsetAsync(s->counts, 0, 6 * sizeof(unsigned long long), s->stream);
    if (e == hipSuccess) e = kf::launch_raycast(a, s->cfg.storage, s->st, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(counts_host, d, np * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    const hipError_t ef = hipFree(d);
    KF_HIP(e);
    KF_HIP(ef);
    return KF_OK;
}

int kf_raycast_wave_trace(kf_session *s, uint32_t *records_host, int *n_records)
{
    if (!s || !records_host || !n_records) return fail(KF_ERR_INVALID, "NULL argument");
    kf::RaycastArgs a = raycast_args(s, 1);
    // the hybrid launch has 4 x heavy_max blocks before the one-lane grid
    const size_t n = ((size_t)a.blocks + (a.heavy_max > 0 && a.tile_cost ? 4 * (size_t)a.heavy_max : 0)) * 4;
    if (n > 4 * 4096) return fail(KF_ERR_INVALID, "too many wavefronts for the trace");
    uint32_t *d = nullptr;
    KF_HIP(hipMalloc(&d, n * 8 * sizeof(uint32_t)));
    a.wave_trace = d;
    hipError_t e = hipMemsetAsync(d, 0, n * 8 * sizeof(uint32_t), s->stream);
    if (e == hipSuccess) e = kf::launch_raycast(a, s->cfg.storage, s->st, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(records_host, d, n * 8 * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    const hipError_t ef = hipFree(d);
    KF_HIP(e);
    KF_HIP(ef);
    *n_records = (int)n;
    return KF_OK;
}

int kf_count_raycast_traffic(kf_session *s, uint64_t *voxels_read, uint64_t *flag_bytes_read)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    const int *N = s->cfg.volume_size;
    const size_t nvox = (size_t)N[0] * N[1] * N[2];
    const size_t nflag = (size_t)s->NB[0] * s->NB[1] * s->NB[2] + (size_t)s->NC[0] * s->NC[1] * s->NC[2];
    const size_t words = (nvox + nflag + 31) / 32, whole = nvox / 32;
    uint32_t *bits = nullptr;
    KF_HIP(hipMalloc(&bits, words * sizeof(uint32_t)));
    kf::RaycastArgs a = raycast_args(s, 1);
    a.touch = bits;
    unsigned long long h = 0;
    std::vector<uint32_t> tail(words - whole);
    // the words holding only voxel bits are counted on the device, the rest
    // (the last partial voxel word and the flag bits) on the host
    hipError_t e = hipMemsetAsync(bits, 0, words * sizeof(uint32_t), s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->counts, 0, sizeof(unsigned long long), s->stream);
    if (e == hipSuccess) e = kf::launch_raycast_touch(a, s->cfg.storage, s->st, whole, s->counts, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, s->counts, sizeof(h), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(tail.data(), bits + whole, tail.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    const hipError_t ef = hipFree(bits);
    KF_HIP(e);
    KF_HIP(ef);
    uint64_t vox = h, flags = 0;
    for (size_t k = 0; k < tail.size(); ++k)
        for (int b = 0; b < 32; ++b)
            if ((tail[k] >> b) & 1u) {
                if ((whole + k) * 32 + (size_t)b < nvox)
                    ++vox;
                else
                    ++flags;
            }
    if (voxels_read) *voxels_read = vox;
    if (flag_bytes_read) *flag_bytes_read = flags;
    return KF_OK;
}

int kf_count_integrate_traffic(kf_session *s, uint64_t *voxels_read, uint64_t *voxels_written)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    KF_HIP(hipMemsetAsync(s->counts, 0, 2 * sizeof(unsigned long long), s->stream));
    kf::IntegrateArgs a = integrate_args(s, 0);
    KF_HIP(kf::launch_count_integrate(a, s->st, s->counts, s->stream));
    unsigned long long h[2];
    KF_HIP(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    if (voxels_read) *voxels_read = h[0];
    if (voxels_written) *voxels_written = h[1];
    return KF_OK;
}

int kf_count_integrate_fast_traffic(kf_session *s, uint64_t *voxels_read, uint64_t *voxels_written)
{
    if (!s) return fail(KF_ERR_INVALID, "session is NULL");
    KF_HIP(hipMemsetAsync(s->counts, 0, 2 * sizeof(unsigned long long), s->stream));
    kf::IntegrateArgs a = integrate_args(s, 0);
    KF_HIP(kf::launch_count_integrate_fast(a, s->cfg.storage, s->st, s->counts, s->stream));
    unsigned long long h[2];
    KF_HIP(hipMemcpyAsync(h, s->counts, sizeof(h), hipMemcpyDeviceToHost, s->stream));
    KF_HIP(hipStreamSynchronize(s->stream));
    if (voxels_read) *voxels_read = h[0];
    if (voxels_written) *voxels_written = h[1];
    return KF_OK;
}

// ---------------------------------------------------------------------------
// Multi-session batches
// ---------------------------------------------------------------------------
}  // extern "C"

struct kf_batch {
    std::vector<kf_session *> sessions;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    kf::BatchSession *d_sess = nullptr;
    int *d_prefix = nullptr;  // integrate and coarse-cell block prefixes: 2 x (kMaxBatch + 1)
    int n = 0;
    int storage = 0, fast = 0, off32 = 1;
    int pyr_blocks = 0, ray_blocks = 0, nblk[3] = {0, 0, 0};
    int ray_lanes = 1;
    int ray_dim_pow2 = 1;  // every member's volume dimensions are powers of two
    int total_int = 0, total_coarse = 0;
    std::vector<long> frames;
    // Membership changes upload new tables with stream-ordered copies from a
    // ring of pinned staging slots, so set_sessions never blocks the host on
    // queued frames: a slot is reused only after its previous copy completed.
    static constexpr int kSlots = 4;
    static constexpr size_t kSlotBytes = sizeof(kf::BatchSession) * kf::kMaxBatch + sizeof(int) * 2 * (kf::kMaxBatch + 1);
    unsigned char *h_stage[kSlots] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t slot_ev[kSlots] = {nullptr, nullptr, nullptr, nullptr};
    bool slot_pending[kSlots] = {false, false, false, false};
    int next_slot = 0;
    hipEvent_t join_ev = nullptr;  // orders the members' own streams before the batch stream
    // Tick records (depth pointers, active mask): a ring of device slots, each
    // filled by a stream-ordered copy from its own pinned host slot; a host
    // slot is rewritten only after the tick that read it completed.
    static constexpr int kTicks = 4;
    kf::BatchTick *d_ticks = nullptr;
    kf::BatchTick *h_ticks = nullptr;
    hipEvent_t tick_ev[kTicks] = {nullptr, nullptr, nullptr, nullptr};
    bool tick_pending[kTicks] = {false, false, false, false};
    int next_tick = 0;
    // Track() of all members as one persistent launch (icp_per blocks per
    // session), or one launch per ICP iteration
    int icp_persistent = 1, icp_per = 0;
    // each ring slot's tick captured as a hipGraph (copy + launches), rebuilt
    // after a membership change (version); off by default: the direct
    // launches are queued ahead of the GPU already and measured as fast
    int use_graph = 0;
    bool graph_failed = false;
    unsigned version = 0;
    hipGraph_t graph[kTicks] = {nullptr, nullptr, nullptr, nullptr};
    hipGraphExec_t gexec[kTicks] = {nullptr, nullptr, nullptr, nullptr};
    unsigned graph_version[kTicks] = {0, 0, 0, 0};
    int cus = 0;
    int iters[3] = {0, 0, 0};
    int int_rows = 4;
};

namespace {

void drop_batch_graphs(kf_batch *b)
{
    for (int k = 0; k < kf_batch::kTicks; ++k) {
        if (b->gexec[k]) (void)hipGraphExecDestroy(b->gexec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->gexec[k] = nullptr;
        b->graph[k] = nullptr;
    }
}

void release_batch(kf_batch *b)
{
    drop_batch_graphs(b);
    for (int k = 0; k < kf_batch::kTicks; ++k)
        if (b->tick_ev[k]) (void)hipEventDestroy(b->tick_ev[k]);
    if (b->d_ticks) (void)hipFree(b->d_ticks);
    if (b->h_ticks) (void)hipHostFree(b->h_ticks);
    if (b->d_sess) (void)hipFree(b->d_sess);
    if (b->d_prefix) (void)hipFree(b->d_prefix);
    for (int k = 0; k < kf_batch::kSlots; ++k) {
        if (b->h_stage[k]) (void)hipHostFree(b->h_stage[k]);
        if (b->slot_ev[k]) (void)hipEventDestroy(b->slot_ev[k]);
    }
    if (b->join_ev) (void)hipEventDestroy(b->join_ev);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int batch_fill(kf_batch *b, kf_session *const *sessions, int n)
{
    if (n < 1 || n > kf::kMaxBatch) return fail(KF_ERR_INVALID, "a batch holds 1..64 sessions");
    for (int i = 0; i < n; ++i)
        if (!sessions[i]) return fail(KF_ERR_INVALID, "NULL session in batch");
    const kf_session *s0 = sessions[0];
    int iters = 0;
    for (int l = 0; l < 3; ++l) iters += s0->cfg.iterations[l];
    if (iters == 0) return fail(KF_ERR_INVALID, "batched sessions need at least one ICP iteration");
    for (int i = 0; i < n; ++i) {
        const kf_session *si = sessions[i];
        if (si->cfg.width != s0->cfg.width || si->cfg.height != s0->cfg.height || si->cfg.storage != s0->cfg.storage ||
            si->fast_math != s0->fast_math || si->icp_rows_per_block != s0->icp_rows_per_block ||
            si->device != s0->device ||
            si->icp_all_groups != s0->icp_all_groups || si->cfg.reduce_rule != s0->cfg.reduce_rule ||
            si->icp_solo_rows != s0->icp_solo_rows ||
            std::memcmp(si->cfg.iterations, s0->cfg.iterations, sizeof(s0->cfg.iterations)) != 0)
            return fail(KF_ERR_INVALID, "batched sessions must share depth size, ICP schedule, storage and options");
        if (si->track_mode != KF_TRACK_POINT_TO_PLANE)
            return fail(KF_ERR_INVALID, "batched sessions track point-to-plane only (KF_OPT_TRACK_MODE = 1)");
        if (si->icp_ref_order)
            return fail(KF_ERR_INVALID, "reference-order ICP sums (KF_OPT_ICP_REFERENCE_ORDER) are single-session only");
        for (int j = 0; j < i; ++j)
            if (sessions[j] == si) return fail(KF_ERR_INVALID, "a session appears twice in the batch");
    }
    // work the members queued on their own streams runs before the next batch launch
    for (int i = 0; i < n; ++i) {
        if (sessions[i]->stream == b->stream) continue;
        KF_HIP(hipEventRecord(b->join_ev, sessions[i]->stream));
        KF_HIP(hipStreamWaitEvent(b->stream, b->join_ev, 0));
    }
    const int slot = b->next_slot;
    if (b->slot_pending[slot]) KF_HIP(hipEventSynchronize(b->slot_ev[slot]));
    b->slot_pending[slot] = false;
    kf::BatchSession *h = reinterpret_cast<kf::BatchSession *>(b->h_stage[slot]);
    int *ip = reinterpret_cast<int *>(b->h_stage[slot] + sizeof(kf::BatchSession) * kf::kMaxBatch);
    int *cp = ip + kf::kMaxBatch + 1;
    std::memset(ip, 0, sizeof(int) * 2 * (kf::kMaxBatch + 1));
    b->off32 = 1;
    for (int i = 0; i < n; ++i) {
        kf_session *si = sessions[i];
        kf::BatchSession &bs = h[i];
        bs.pyr = pyramid_args(si);
        for (int l = 0; l < 3; ++l) bs.icp[l] = icp_args(si, l, 0, 0);
        bs.trk = kf::IcpTrackArgs{};
        for (int l = 0; l < 3; ++l) {
            bs.trk.lv[l] = bs.icp[l];
            bs.trk.iters[l] = si->cfg.iterations[l];
        }
        bs.trk.frame = 1;
        bs.trk.rows[0] = si->icp_rows;
        bs.trk.rows[1] = si->icp_rows + (size_t)(si->cfg.height + 8) * 32;
        bs.trk.spin_limit = 1u << 22;
        // launch-wide shapes (integrate block rows, raycast lanes and XCD
        // clusters) are the first member's; they do not change any result
        bs.integ = integrate_args(si, 1, 1048576);
        bs.integ.rows = s0->int_rows;
        bs.ray = raycast_args(si, s0->ray_lanes, s0->xcd_group);
        bs.brick = brick_args(si);
        bs.st = si->st;
        kf::integrate_grid_dims(bs.integ, bs.int_grid);
        ip[i + 1] = ip[i] + bs.int_grid[0] * bs.int_grid[1] * bs.int_grid[2];
        cp[i + 1] = cp[i] + (int)si->ncoarse;
        if (!kf::integrate_planes_fit_32(bs.integ, si->cfg.storage)) b->off32 = 0;
    }
    // stream order: launches queued earlier still read the previous tables
    KF_HIP(hipMemcpyAsync(b->d_sess, h, sizeof(kf::BatchSession) * n, hipMemcpyHostToDevice, b->stream));
    KF_HIP(hipMemcpyAsync(b->d_prefix, ip, sizeof(int) * 2 * (kf::kMaxBatch + 1), hipMemcpyHostToDevice, b->stream));
    KF_HIP(hipEventRecord(b->slot_ev[slot], b->stream));
    b->slot_pending[slot] = true;
    b->next_slot = (slot + 1) % kf_batch::kSlots;
    b->sessions.assign(sessions, sessions + n);
    b->frames.assign(n, 0);
    for (int i = 0; i < n; ++i) b->frames[i] = sessions[i]->frames;
    b->n = n;
    b->storage = s0->cfg.storage;
    b->fast = s0->fast_math;
    b->pyr_blocks = h[0].pyr.blocks_total;
    b->ray_blocks = h[0].ray.blocks;
    b->ray_lanes = h[0].ray.lanes;
    b->ray_dim_pow2 = 1;
    for (int i = 0; i < n; ++i) b->ray_dim_pow2 = b->ray_dim_pow2 && h[i].ray.dim_pow2;
    for (int l = 0; l < 3; ++l) b->nblk[l] = (h[0].icp[l].group_only >= 0 ? 1 : 8) * h[0].icp[l].rows_step;
    b->total_int = ip[n];
    b->total_coarse = cp[n];
    for (int l = 0; l < 3; ++l) b->iters[l] = s0->cfg.iterations[l];
    b->int_rows = s0->int_rows;
    // persistent Track(): every member's blocks resident at once, at most one
    // block per compute unit, no more blocks per session than the largest
    // level has partial rows
    int maxnb = 0;
    bool pers = true;
    for (int l = 0; l < 3; ++l)
        if (h[0].icp[l].rows_step > 0 && b->iters[l] > 0) maxnb = std::max(maxnb, kf::icp_blocks(h[0].icp[l]));
    for (int i = 0; i < n; ++i) pers = pers && sessions[i]->icp_persistent && !sessions[i]->debug;
    b->icp_per = pers ? std::min(maxnb, b->cus / n) : 0;
    ++b->version;  // captured ticks hold the old launch shapes
    return KF_OK;
}

// The launches of one tick reading ring slot k: the record copy, then the
// frame block of every member (kinect.cpp:654-695).
int enqueue_tick(kf_batch *b, int k)
{
    const kf::BatchTick *tp = b->d_ticks + k;
    kf::BatchTable tb{};
    tb.sess = b->d_sess;
    tb.int_prefix = b->d_prefix;
    tb.coa_prefix = b->d_prefix + (kf::kMaxBatch + 1);
    tb.n = b->n;
    KF_HIP(hipMemcpyAsync(b->d_ticks + k, b->h_ticks + k, sizeof(kf::BatchTick), hipMemcpyHostToDevice, b->stream));
    hipError_t e = kf::launch_batch_pyramid(tb, tp, b->pyr_blocks, b->stream);
    if (b->icp_persistent && b->icp_per > 0) {
        if (e == hipSuccess) e = kf::launch_batch_icp_track(tb, tp, b->icp_per, b->stream);
    } else {
        int last_level = -1;
        for (int level = 0; level <= 2 && last_level < 0; ++level)
            if (b->iters[level] > 0) last_level = level;
        bool first_any = true;
        for (int level = 2; level >= 0 && e == hipSuccess; --level)
            for (int it = 0; it < b->iters[level] && e == hipSuccess; ++it) {
                const int fin = level == last_level && it == b->iters[level] - 1;
                e = kf::launch_batch_icp(tb, tp, level, it == 0, first_any ? 1 : 0, fin, b->nblk[level], b->stream);
                first_any = false;
            }
    }
    if (e == hipSuccess)
        e = kf::launch_batch_integrate(tb, tp, b->storage, b->fast, b->off32, b->total_int, b->int_rows, b->stream);
    if (e == hipSuccess) e = kf::launch_batch_bricks(tb, tp, b->total_coarse, b->stream);
    if (e == hipSuccess) e = kf::launch_batch_raycast(tb, tp, b->storage, b->ray_lanes, b->ray_blocks, b->ray_dim_pow2,
                                                            b->stream);
    if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(e));
    return KF_OK;
}

// Replays ring slot k's captured tick, capturing it first when the
// membership changed since; falls back to direct launches when capture is
// not possible on the batch stream.
int run_tick(kf_batch *b, int k)
{
    if (!b->use_graph || b->graph_failed) return enqueue_tick(b, k);
    if (!b->gexec[k] || b->graph_version[k] != b->version) {
        if (b->gexec[k]) (void)hipGraphExecDestroy(b->gexec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->gexec[k] = nullptr;
        b->graph[k] = nullptr;
        if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            b->graph_failed = true;
            return enqueue_tick(b, k);
        }
        const int rc = enqueue_tick(b, k);
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        if (rc != KF_OK || e2 != hipSuccess || !g ||
            hipGraphInstantiate(&b->gexec[k], g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            b->gexec[k] = nullptr;
            (void)hipGetLastError();
            b->graph_failed = true;
            return rc != KF_OK ? rc : enqueue_tick(b, k);
        }
        b->graph[k] = g;
        b->graph_version[k] = b->version;
    }
    KF_HIP(hipGraphLaunch(b->gexec[k], b->stream));
    return KF_OK;
}

}  // namespace

extern "C" {

int kf_batch_create(kf_session *const *sessions, int n, void *stream, kf_batch **out)
{
    if (!out) return fail(KF_ERR_INVALID, "out is NULL");
    *out = nullptr;
    kf_batch *b = new (std::nothrow) kf_batch();
    if (!b) return fail(KF_ERR_OOM, "host allocation failed");
    auto cleanup = [&](int code) {
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        release_batch(b);
        return code;
    };
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(KF_ERR_HIP, "hipStreamCreate failed"));
        b->own_stream = true;
    }
    if (hipMalloc((void **)&b->d_sess, sizeof(kf::BatchSession) * kf::kMaxBatch) != hipSuccess ||
        hipMalloc((void **)&b->d_prefix, sizeof(int) * 2 * (kf::kMaxBatch + 1)) != hipSuccess)
        return cleanup(fail(KF_ERR_OOM, "batch tables"));
    for (int k = 0; k < kf_batch::kSlots; ++k)
        if (hipHostMalloc((void **)&b->h_stage[k], kf_batch::kSlotBytes, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&b->slot_ev[k], hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(KF_ERR_OOM, "batch staging"));
    if (hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(KF_ERR_HIP, "hipEventCreate failed"));
    if (hipMalloc((void **)&b->d_ticks, sizeof(kf::BatchTick) * kf_batch::kTicks) != hipSuccess ||
        hipHostMalloc((void **)&b->h_ticks, sizeof(kf::BatchTick) * kf_batch::kTicks, hipHostMallocDefault) !=
            hipSuccess)
        return cleanup(fail(KF_ERR_OOM, "batch tick records"));
    for (int k = 0; k < kf_batch::kTicks; ++k)
        if (hipEventCreateWithFlags(&b->tick_ev[k], hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(KF_ERR_HIP, "hipEventCreate failed"));
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return cleanup(fail(KF_ERR_HIP, "cannot query the compute unit count"));
    }
    int rc = batch_fill(b, sessions, n);
    if (rc) return cleanup(rc);
    *out = b;
    return KF_OK;
}

int kf_batch_set_sessions(kf_batch *b, kf_session *const *sessions, int n)
{
    if (!b) return fail(KF_ERR_INVALID, "batch is NULL");
    return batch_fill(b, sessions, n);
}

int kf_batch_destroy(kf_batch *b)
{
    if (!b) return KF_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    release_batch(b);
    return KF_OK;
}

int kf_batch_synchronize(kf_batch *b)
{
    if (!b) return fail(KF_ERR_INVALID, "batch is NULL");
    KF_HIP(hipStreamSynchronize(b->stream));
    return KF_OK;
}

int kf_batch_get_stream(kf_batch *b, void **stream)
{
    if (!b || !stream) return fail(KF_ERR_INVALID, "NULL argument");
    *stream = (void *)b->stream;
    return KF_OK;
}

int kf_batch_set_option(kf_batch *b, int option, int value)
{
    if (!b) return fail(KF_ERR_INVALID, "batch is NULL");
    switch (option) {
    case KF_OPT_GRAPH:
        b->use_graph = value ? 1 : 0;
        b->graph_failed = false;
        break;
    case KF_OPT_ICP_PERSISTENT: b->icp_persistent = value ? 1 : 0; break;
    default: return fail(KF_ERR_INVALID, "unknown batch option");
    }
    drop_batch_graphs(b);
    return KF_OK;
}

int kf_batch_get_option(kf_batch *b, int option, int *value)
{
    if (!b || !value) return fail(KF_ERR_INVALID, "NULL argument");
    switch (option) {
    case KF_OPT_GRAPH: *value = b->use_graph && !b->graph_failed; break;
    case KF_OPT_ICP_PERSISTENT: *value = b->icp_persistent && b->icp_per > 0; break;
    default: return fail(KF_ERR_INVALID, "unknown batch option");
    }
    return KF_OK;
}

int kf_batch_frame_u16_dev(kf_batch *b, const uint16_t *const *depth_mm_dev, const int *active)
{
    if (!b || !depth_mm_dev) return fail(KF_ERR_INVALID, "NULL argument");
    const int k = b->next_tick;
    if (b->tick_pending[k]) KF_HIP(hipEventSynchronize(b->tick_ev[k]));
    b->tick_pending[k] = false;
    kf::BatchTick &t = b->h_ticks[k];
    std::memset(&t, 0, sizeof(t));
    int nact = 0;
    for (int i = 0; i < b->n; ++i) {
        const bool on = active ? active[i] != 0 : true;
        if (!on) continue;
        if (!depth_mm_dev[i]) return fail(KF_ERR_INVALID, "active session without depth");
        t.active |= 1ull << i;
        t.depth[i] = depth_mm_dev[i];
        ++nact;
    }
    if (!nact) return KF_OK;
    int rc = run_tick(b, k);
    if (rc) return rc;
    KF_HIP(hipEventRecord(b->tick_ev[k], b->stream));
    b->tick_pending[k] = true;
    b->next_tick = (k + 1) % kf_batch::kTicks;
    hipError_t e = hipSuccess;
    // per-session exact brick rebuild on the same schedule as kf_frame_*
    for (int i = 0; i < b->n; ++i) {
        if (!((t.active >> i) & 1ull)) continue;
        kf_session *si = b->sessions[i];
        si->frames = ++b->frames[i];
        if (si->rebuild_period > 0 && si->frames % si->rebuild_period == 0) {
            kf::BrickArgs ba = brick_args(si);
            ba.incremental = 1;
            e = kf::launch_brick_rebuild(ba, si->cfg.storage, b->stream);
            if (e == hipSuccess) e = kf::launch_brick_dilate(ba, b->stream);
            if (e != hipSuccess) return fail(KF_ERR_HIP, std::string("brick rebuild: ") + hipGetErrorString(e));
        }
    }
    return KF_OK;
}

}  // extern "C"
