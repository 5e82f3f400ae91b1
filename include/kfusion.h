/*
 * kfusion.h -- C ABI of the MI355X (gfx950) KinectFusion frame loop.
 *
 * Drop-in boundary for the hot path of philipNoonan/oglFusionKinectV2: the
 * GL textures / SSBOs owned by `class gFusion` (include/gFusion.h:124-276)
 * become device buffers owned by an opaque session; every GLSL dispatch of the
 * frame loop becomes a hand-written HIP kernel launched on the session's
 * stream.  Conventions:
 *   - every entry point returns an int status (KF_OK = 0, negative = error);
 *     kf_last_error() gives a message for the calling thread;
 *   - pointers are HOST pointers unless the name ends in _dev;
 *   - matrices are 16 floats, column-major (glm layout: m[c*4 + r]);
 *   - "pose" is camera->world (gFusion::m_pose);
 *   - a session is driven by one thread at a time; sessions are independent;
 *   - no exceptions cross the ABI.
 */
#ifndef KFUSION_H
#define KFUSION_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define KF_OK 0
#define KF_ERR_INVALID (-1) /* bad argument or configuration              */
#define KF_ERR_HIP (-2)     /* a HIP runtime call failed                   */
#define KF_ERR_OOM (-3)     /* device allocation failed                    */
#define KF_ERR_BUDGET (-4)  /* admission refused: HBM budget exceeded      */
#define KF_ERR_STATE (-5)   /* call not valid in the current session state */

/* ---- TSDF storage (voxel = value + weight, stored as two planes) ---------- */
#define KF_STORAGE_RG16I 0 /* int16 value*32766 + int16 weight (reference RG16I, 4 B) */
#define KF_STORAGE_F32 1   /* fp32 value + fp32 weight (8 B)                            */
#define KF_STORAGE_F16 2   /* fp16 value + fp16 weight (4 B), fp32 update arithmetic   */

/* ---- combine rule for the 8x32 ICP partials (gFusion.cpp:1143-1149) ------- */
#define KF_RULE_LAST_ROW 0 /* reference behaviour: only work group 7 survives */
#define KF_RULE_SUM_ALL 1  /* sum the 8 rows                                 */

/* ---- map selectors for kf_read_map / kf_write_map ------------------------- */
#define KF_MAP_DEPTH 0      /* float mm, level l                    */
#define KF_MAP_VERTEX 1     /* float4 camera-space vertex, level l  */
#define KF_MAP_NORMAL 2     /* float4 normal (or (2,0,0,0)), level l */
#define KF_MAP_REF_VERTEX 3 /* float4 raycast world vertex, level 0  */
#define KF_MAP_REF_NORMAL 4 /* float4 raycast world normal, level 0  */
#define KF_MAP_INTENSITY 5  /* float4 intensity projection, level 0 (kf_intensity_projection) */

/* ---- integrate gating ------------------------------------------------------ */
#define KF_INTEGRATE_ALWAYS 0 /* gFusion::integrate() */
#define KF_INTEGRATE_GATED 1  /* only if the last kf_frame/kf_track_frame decided so (device flag) */

/* gFusionConfig (include/gFusion.h:42-82) plus the camera and depth size. */
typedef struct kf_config {
    int volume_size[3];     /* voxels                        (volumeSize)       */
    float volume_dim[3];    /* metres                        (volumeDimensions) */
    float mu;               /* truncation distance, metres                      */
    float max_weight;
    float near_plane, far_plane;
    float dist_threshold, normal_threshold;
    int iterations[3];      /* ICP iterations per level, 0 = finest             */
    float cam[4];           /* fx, fy, cx, cy of the depth camera (camPamsDepth) */
    int width, height;      /* depth image size (512 x 424)                     */
    int storage;            /* KF_STORAGE_*                                     */
    int reduce_rule;        /* KF_RULE_*                                        */
} kf_config;

typedef struct kf_session kf_session;

/* Fills the defaults of gFusionConfig() (include/gFusion.h:55-70), the Kinect
 * v2 intrinsics used by the synthetic sequences and a 512x424 depth image. */
int kf_config_default(kf_config *cfg);

/* Device bytes a session with this configuration holds (admission control). */
int kf_session_bytes(const kf_config *cfg, size_t *bytes);

/* Creates a session: allocates the device buffers that initTextures /
 * initVolume / allocateBuffers create (gFusion.cpp:200-330).  `stream` is a
 * hipStream_t (NULL: the session creates its own).  The volume is zeroed and
 * the pose set to identity. */
int kf_session_create(const kf_config *cfg, void *stream, kf_session **out);
int kf_session_destroy(kf_session *s);
int kf_get_config(const kf_session *s, kf_config *cfg);

/* Reset(pose, false) + raycast(), as the UI does (gFusion.cpp:226-287,
 * kinect.cpp:446-477): zero the volume, set the pose, raycast the empty
 * volume, and arm the 4-frame integrate window of the frame loop. */
int kf_reset(kf_session *s, const float pose[16]);
int kf_reset_volume(kf_session *s);              /* resetVolume (gFusion.cpp:243) */
int kf_set_pose(kf_session *s, const float pose[16]); /* setPose (gFusion.h:184)  */
int kf_get_pose(kf_session *s, float pose[16]);  /* getPose (gFusion.h:180), syncs */

/* depthToVertex(float*) (gFusion.cpp:496): upload one 512x424 depth frame in
 * millimetres.  _f32 takes the reference's float mm array, _u16 the sensor's
 * uint16 mm, _dev variants read a device buffer already in HBM.  The upload
 * also builds the pyramid (mip + vertex + normal maps, gFusion.cpp:496-547). */
int kf_depth_to_vertex_f32(kf_session *s, const float *depth_mm);
int kf_depth_to_vertex_u16(kf_session *s, const uint16_t *depth_mm);
int kf_depth_to_vertex_u16_dev(kf_session *s, const uint16_t *depth_mm_dev);
int kf_depth_to_vertex_f32_dev(kf_session *s, const float *depth_mm_dev);

/* Track() (gFusion.cpp:648-740): projective point-to-plane ICP over the 3
 * levels, coarse to fine, entirely on the device (no host round trip per
 * iteration).  If `tracked` is non-NULL the call synchronises and returns the
 * accept/reject decision (0 < AE <= 0.5); the pose is restored on reject. */
int kf_track(kf_session *s, int *tracked);

/* The tracking step as the frame loop runs it: Track() and / or TrackSDF()
 * (KF_OPT_TRACK_MODE), then the frame-loop gate of kinect.cpp:677-691
 * (integrate if tracked or within the reset window; frame counter) -- all on
 * the device, no sync. */
int kf_track_frame(kf_session *s);

/* One ICP iteration at `level` (track(level) + reduce(level) + getReduction +
 * solve + pose update, gFusion.cpp:669-708).  `first` = 1 starts a level
 * (clears the early-exit flag); `save_pose` = 1 records the pose to restore. */
int kf_icp_iteration(kf_session *s, int level, int first, int save_pose);
/* The accept/reject step of Track() after the last iteration (gFusion.cpp:715-723). */
int kf_icp_finish(kf_session *s);

/* TrackSDF() (gFusion.cpp:835-1038, trackSDF.cs, reduceSDF.cs): point-to-
 * volume Gauss-Newton tracking of the level vertex maps against the TSDF,
 * coarse to fine, starting from m_pose_eig (set by kf_reset to the reset
 * pose's translation with identity rotation, and moved only by a successful
 * TrackSDF).  One persistent launch, no host round trip.  On success the pose
 * becomes exp(result) * m_pose_eig; otherwise it is unchanged.  If `tracked`
 * is non-NULL the call synchronises and returns the result. */
int kf_track_sdf(kf_session *s, int *tracked);

/* recoverPose() (gFusion.cpp:1616-1653): for each pose-library entry in order,
 * set the pose to it and run TrackSDF, stopping at the first success.  The
 * library holds the poses of successful kf_track / frame-loop Track() calls
 * that lie 0.2 m or more from every earlier entry (updatePoseFinder,
 * gFusion.cpp:1570-1597), at most KF_POSE_LIBRARY_MAX of them; kf_reset
 * clears it.  If `found` is non-NULL the call synchronises and returns it. */
int kf_recover_pose(kf_session *s, int *found);
#define KF_POSE_LIBRARY_MAX 256
/* The library poses (16 floats each, column-major) into `poses` (room for
 * `capacity` poses, may be NULL) and their number into `count`. */
int kf_read_pose_library(kf_session *s, float *poses, int capacity, int *count);
/* getCumTwist (gFusion.h:260-265): the sum of the twists of successful TrackSDF calls. */
int kf_get_cum_twist(kf_session *s, float twist[6]);
/* TrackSDF state (any pointer may be NULL): m_pose_eig, m_cumTwist, the 8x32
 * partials of the last iteration (reduceSDF.cs layout) and its alignment
 * energy, iterations run by the last TrackSDF / recoverPose launch, and a log
 * of 64 rows x 10 doubles (level, iteration, AE, |step|, result[6]).  Returns
 * KF_ERR_STATE if a launch was aborted because its grid barrier timed out. */
int kf_read_sdf_state(kf_session *s, float pose_eig[16], float cum_twist[6], float partials[256], float *ae,
                      int *iterations_run, double *log);
/* Overwrites m_pose_eig (no counterpart in the reference, which sets it only
 * in resetPose and TrackSDF; used to load a known tracking state). */
int kf_write_pose_eig(kf_session *s, const float pose_eig[16]);

/* intensityProjection() (gFusion.cpp:1368-1409, mip.cs): for every pixel the
 * sum of the positive TSDF samples along its ray, one every 0.75 mu from the
 * box / near-plane entry to min(box exit, far plane), as (sum, 0, 0, 1) into
 * the KF_MAP_INTENSITY image (kf_read_map).  The first call allocates the
 * image (W x H x 16 B) and a flag buffer (one byte per 8^3 brick and per 32^3
 * cell).  With KF_OPT_SKIP_EMPTY, samples whose trilinear corners lie only in
 * bricks holding (0, 0) are not evaluated (they are 0 and add nothing: the
 * image is the same bit for bit). */
int kf_intensity_projection(kf_session *s);

/* integrate() (gFusion.cpp:1234-1287).  mode = KF_INTEGRATE_ALWAYS or _GATED. */
int kf_integrate(kf_session *s, int mode);
/* raycast() (gFusion.cpp:1289-1366): refreshes the reference vertex/normal maps. */
int kf_raycast(kf_session *s);

/* The frame block of kinect.cpp:654-695 with integration on: pyramid, Track
 * and / or TrackSDF (KF_OPT_TRACK_MODE), integrate if tracked or within the
 * reset window, recoverPose if the tracking failed (KF_OPT_RECOVER), raycast.
 * No host synchronisation; the tracking result stays on the device
 * (kf_read_status). */
int kf_frame_u16_dev(kf_session *s, const uint16_t *depth_mm_dev);
int kf_frame_f32(kf_session *s, const float *depth_mm);

/* Read-back surface (gFusion.h:180-274).  All synchronise the session stream. */
int kf_read_map(kf_session *s, int which, int level, void *host, size_t bytes);
int kf_write_map(kf_session *s, int which, int level, const void *host, size_t bytes);
/* TSDF planes in the session's storage type, x fastest then y then z.  Either
 * pointer may be NULL. */
int kf_read_tsdf(kf_session *s, void *value, void *weight);
int kf_write_tsdf(kf_session *s, const void *value, const void *weight);
/* The 8x32 partial sums of the last ICP iteration (reduce.cs output layout),
 * its alignment energy, and the frame state: tracked flag, ICP iterations run,
 * frame counter and reset window flag. */
int kf_read_reduction(kf_session *s, float partials[256], float *alignment_energy);
int kf_read_status(kf_session *s, int *tracked, int *iterations_run, int *counter, int *reset);

/* The 6-vector of the last ICP solve (gFusion.cpp:696-697: translation then
 * the three angles of the pose update).  No counterpart in the reference (a
 * test aid: the pose update can be checked against it exactly). */
int kf_read_icp_solution(kf_session *s, double x[6]);
/* Per-pixel output of the last ICP iteration at `level`
 * ({int result; float error; float J[6]} of track.cs:18-27, as planes).
 * Only recorded when enabled with kf_set_debug(s, 1). */
int kf_set_debug(kf_session *s, int record_track_pixels);
int kf_read_track_pixels(kf_session *s, int level, int32_t *result, float *error, float *J);

/* Tuning options (no counterpart in the reference; all keep its semantics).
 *   KF_OPT_FAST_MATH      integrate uses hardware reciprocal / square root
 *                         (about 1 ulp; GLSL itself allows 2.5 ulp division)
 *                         instead of correctly rounded division.  Default 1;
 *                         0 reproduces the CPU oracle bit for bit.
 *   KF_OPT_SKIP_EMPTY     raycast skips samples in 8^3 bricks that hold no
 *                         negative value (exact: same hits); the intensity
 *                         projection skips samples in bricks of (0, 0)
 *                         (exact: same image).  Default 1.
 *   KF_OPT_GRAPH          kf_frame_* replays a captured hipGraph.  Default 1.
 *   KF_OPT_BRICK_REBUILD_PERIOD  frames between exact brick-flag rebuilds
 *                         (flags are otherwise only set): the raycast's
 *                         occupancy flags, and the zero flags with which
 *                         integrate neither reads nor rewrites bricks that
 *                         hold only (0, 0).  Default 16, 0 = never. */
#define KF_OPT_FAST_MATH 1
#define KF_OPT_SKIP_EMPTY 2
#define KF_OPT_GRAPH 3
#define KF_OPT_BRICK_REBUILD_PERIOD 4
/*   KF_OPT_RAYCAST_XCD_GROUP  raycast tiles per XCD cluster: block b runs on
 *                         XCD b % 8 and receives clusters of this many
 *                         consecutive 16x16 tiles.  Default 4. */
#define KF_OPT_RAYCAST_XCD_GROUP 5
/*   KF_OPT_ICP_ROWS_PER_BLOCK rows of one row group per ICP block.  Default 1. */
#define KF_OPT_ICP_ROWS_PER_BLOCK 6
/*   KF_OPT_STAGE_TIMING   1: frames run without the graph and the pyramid,
 *                         ICP (first iteration .. finish), integrate and
 *                         raycast launches record start / stop events in
 *                         their own dispatch packets; kf_read_stage_times
 *                         returns the last durations.  Default 0. */
#define KF_OPT_STAGE_TIMING 7
/*   KF_OPT_ICP_ALL_GROUPS 1: every ICP iteration sums all 8 row groups of
 *                         reduce.cs.  0 (default): under KF_RULE_LAST_ROW only
 *                         row group 7, the one getReduction reads, is summed
 *                         (same pose, AE and decisions); the other rows of
 *                         kf_read_reduction's partials then hold zero sums.
 *                         Per-pixel recording (kf_set_debug) sums all groups. */
#define KF_OPT_ICP_ALL_GROUPS 10
/*   KF_OPT_RAYCAST_LANES  lanes that march one ray together: 1 (default) or
 *                         4 (the four evaluate 16 consecutive fine-step
 *                         samples per round).  Same results either way. */
#define KF_OPT_RAYCAST_LANES 11
/*   KF_OPT_ICP_PERSISTENT 1: Track() runs as one persistent launch (every
 *                         level and iteration, grid barriers between them);
 *                         0: one launch per iteration.  Same results.
 *                         Default 1. */
#define KF_OPT_ICP_PERSISTENT 12
/*   KF_OPT_ICP_SOLO_ROWS  at a level whose row groups have at most this many
 *                         rows one block takes all rows of a group; when
 *                         only one group is summed (last-row rule) the
 *                         persistent launch runs that level in that block
 *                         alone, with no grid barrier.  0 = off (default:
 *                         one 256-thread block is slower than a barrier). */
#define KF_OPT_ICP_SOLO_ROWS 13
/*   KF_OPT_ICP_REFERENCE_ORDER  1: every ICP iteration records track.cs's
 *                         per-pixel output and forms reduce.cs's sums in
 *                         reduce.cs's own order (8 groups x 112 threads,
 *                         serial FP32 per thread, serial 112-row add,
 *                         reduce.cs:27-127), combines them in FP32
 *                         (gFusion.cpp:1137-1167) and solves by Jacobi
 *                         alone: the partials, AE and pose of the CPU
 *                         restatement bit for bit.  Two launches per
 *                         iteration, single session only (a parity mode).
 *                         0 (default): the fused kernels (same products,
 *                         wave-tree and FP64 group sums). */
#define KF_OPT_ICP_REFERENCE_ORDER 16
/*   KF_OPT_INTEGRATE_ZLEN z voxels one integrate thread walks (a multiple of
 *                         8, up to 4096); 0 (default): enough segments for
 *                         about 2M column segments per launch.  Same results.
 *   KF_OPT_INTEGRATE_ROWS y rows of an integrate block of 64 x rows threads:
 *                         1, 2 or 4 (default).  Same results.
 *   Both, with KF_OPT_INTEGRATE_COLUMNS, KF_OPT_RAYCAST_XCD_GROUP,
 *   KF_OPT_RAYCAST_LANES, KF_OPT_RAYCAST_HYBRID and
 *   KF_OPT_BRICK_REBUILD_PERIOD, are picked per volume size by the autotuner
 *   (oglfusionkinectv2_amd/autotune.py), whose table kf_session_create
 *   applies; kf_set_option overrides it. */
#define KF_OPT_INTEGRATE_ZLEN 17
#define KF_OPT_INTEGRATE_ROWS 18
/*   KF_OPT_POSE_FP64      1: every ICP pose update is composed in FP64
 *                         (delta in double times an FP64 copy of the pose,
 *                         then rounded to the float pose the kernels use),
 *                         the precision variant's SE(3) composition; the
 *                         reference multiplies in float (gFusion.cpp:700-705).
 *                         0 (default): float, as the reference. */
#define KF_OPT_POSE_FP64 19
/*   KF_OPT_INTEGRATE_COLUMNS  1 (default): the fast integrate's per-column
 *                         set-up (inverse pose, the column's position at
 *                         z = 0, its frustum range) runs once per frame in
 *                         one thread per column, not in every z segment's
 *                         thread; 0: per thread.  Same bits either way. */
#define KF_OPT_INTEGRATE_COLUMNS 20
/*   KF_OPT_RAYCAST_HYBRID 1 (default): raycast() marches the tiles whose rays
 *                         took the most loop iterations in the session's
 *                         earlier raycasts with four lanes per ray and every
 *                         other tile with one (one launch; the same bits as
 *                         one lane everywhere); each depth-pyramid launch
 *                         (kf_frame_*, kf_depth_to_vertex_*) picks the tiles
 *                         for the raycasts after it; 0: one lane everywhere.
 *                         Applies with KF_OPT_RAYCAST_LANES 1. */
#define KF_OPT_RAYCAST_HYBRID 21
/*   KF_OPT_RAYCAST_HEAVY_TILES  read-only (kf_get_option; synchronises the
 *                         session stream): the number of 16x16 tiles the
 *                         next hybrid raycast marches with four lanes per
 *                         ray. */
#define KF_OPT_RAYCAST_HEAVY_TILES 22
/*   KF_OPT_TRACK_MODE     tracking of the frame loop (kinect.h:143-144):
 *                         KF_TRACK_POINT_TO_PLANE (Track, default),
 *                         KF_TRACK_POINT_TO_VOLUME (TrackSDF) or both (Track
 *                         then TrackSDF, whose result counts). */
#define KF_OPT_TRACK_MODE 8
#define KF_TRACK_POINT_TO_PLANE 1
#define KF_TRACK_POINT_TO_VOLUME 2
/*   KF_OPT_RECOVER        1: the frame loop calls recoverPose when its
 *                         tracking failed (kinect.cpp:686-689).  Default 1. */
#define KF_OPT_RECOVER 9
int kf_set_option(kf_session *s, int option, int value);
int kf_get_option(kf_session *s, int option, int *value);
/* Durations in ms of the last pyramid, ICP, integrate and raycast stages run
 * with KF_OPT_STAGE_TIMING on (synchronises the session stream).  No
 * counterpart in the reference (its GL timer queries are commented out,
 * gFusion.cpp:729-736, 783-807). */
int kf_read_stage_times(kf_session *s, float ms[4]);
/* Recompute the brick occupancy and zero flags from the volume (done
 * automatically by kf_reset / kf_write_tsdf and periodically by kf_frame_*). */
int kf_rebuild_bricks(kf_session *s);
/* Copies the 8^3 brick flags to the host (synchronises the session stream):
 * occupancy (the brick holds a value that can stop a ray or switch its step:
 * negative or >= 0.79) and zero (the brick holds a voxel other than (0, 0)),
 * NB0 x NB1 x NB2 bytes each (NB_k = ceil(volume_size[k] / 8)), x fastest.
 * Right after a rebuild both are exact; between rebuilds they are supersets
 * of the exact flags.  No counterpart in the reference (a test aid). */
int kf_read_brick_flags(kf_session *s, uint8_t *occupancy_host, uint8_t *zero_host);

/* Synchronise the session stream / get its stream (hipStream_t). */
int kf_synchronize(kf_session *s);
int kf_get_stream(kf_session *s, void **stream);

/* Algorithmic traffic of integrate for the current depth frame and pose:
 * voxels read and voxels written under the reference semantics (no timing). */
int kf_count_integrate_traffic(kf_session *s, uint64_t *voxels_read, uint64_t *voxels_written);

/* The same for the default (KF_OPT_FAST_MATH) integrate as built: voxels it
 * reads and writes for the current depth frame and pose (it skips chunks that
 * stay (0, 0) and stores of (0, 0) over (0, 0)); a walk with no stores.  No
 * counterpart in the reference (a measurement aid). */
int kf_count_integrate_fast_traffic(kf_session *s, uint64_t *voxels_read, uint64_t *voxels_written);

/* Runs raycast() once more with counters: samples evaluated, samples skipped
 * as empty space, empty cells entered, rays that hit, the largest number of
 * samples one ray evaluated, and the largest evaluated + cells of one ray. */
int kf_raycast_stats(kf_session *s, uint64_t stats[6]);

/* The same with per-ray figures: counts_host[y * width + x] = samples the ray
 * evaluated (low 16 bits) | empty cells it entered (high 16 bits), saturated
 * at 65535. */
int kf_raycast_ray_counts(kf_session *s, uint32_t *counts_host);

/* Runs raycast() once more recording, for every wavefront, when it started and
 * ended and where it ran: records_host[8 * k .. 8 * k + 7] for wavefront k =
 * 4 * block + wave = {start (low, high 32 bits), end (low, high) of the
 * 100 MHz GPU clock, the HW_ID register (SIMD, CU, shader engine), the XCC
 * id, the 16x16 pixel tile, 0xAB << 24 | lanes per ray << 16 | the
 * wavefront's largest loop iteration count (empty cells entered + fine-step
 * batches of four samples, as one lane counts them; at most 65535)}; with the
 * hybrid schedule (KF_OPT_RAYCAST_HYBRID) the first 4 x 48 blocks march
 * quarters of the heavy tiles four lanes per ray and the one-lane blocks of
 * those tiles hold no pixel.  Wavefronts that hold no pixel leave their
 * record zero.  *n_records receives the number of records (4 x blocks);
 * records_host must hold 8 x that many words (at most 8 x 4 x 4096).  Results
 * identical to raycast().  No counterpart in the reference (a measurement
 * aid). */
int kf_raycast_wave_trace(kf_session *s, uint32_t *records_host, int *n_records);

/* Algorithmic traffic of raycast() for the current volume and pose: the
 * number of distinct voxels whose value the march and the normals read, and
 * of distinct occupancy-flag bytes it read (no timing; the maps written are
 * identical to raycast()'s). */
int kf_count_raycast_traffic(kf_session *s, uint64_t *voxels_read, uint64_t *flag_bytes_read);

/* ---- input adapters --------------------------------------------------------
 * camPams of kinect.h (the depth camera's intrinsics and distortion). */
typedef struct kf_cam_params {
    float fx, fy, ppx, ppy;
    float k1, k2, p1, p2, k3;
} kf_cam_params;
/* loadCalibration (kinect.cpp:78-161): an OpenCV-style XML file (the <data>
 * of Camera_Matrix, 3x3 row-major, and of Distortion_Coefficients, 5 values)
 * or a JSON file ("camera_matrix" 3x3, "dist_coeff" 5 values), chosen by the
 * extension.  The reference's JSON branch always opens
 * resources/infrared.json; here the file named is read.  KF_ERR_INVALID for
 * another extension or a malformed file (the reference returns 0 / a tinyxml2
 * error code).  The frame loop then sets the camera to (fx, fx, ppx, ppy)
 * (kinect.cpp:176-177). */
int kf_load_calibration(const char *path, kf_cam_params *out);

/* ---- surface export (MCubes, mcubes.cpp / histoPyramids.cs) ----------------
 * Marching cubes over the session's TSDF as "save stl" runs it
 * (kinect.cpp:487-504: gridSize = volumeSize, iso level 0): cubes within 10
 * voxels of the volume's border and cubes with no value > 0 among their first
 * seven corners produce nothing (histoPyramids.cs:110-152); triangles are
 * numbered in the order of the reference's histopyramid traversal (Morton
 * order of the cubes, x then z then y fastest per level) and their vertices
 * are (x, y, z, 0) in voxel-index units, interpolated along the cube edges
 * (traverseHistoPyramids.cs:190-250).  The triangle table is built from the
 * cube topology (ambiguous faces separate the below-iso corners), which gives
 * every case the reference table's triangle count and surface polygon. */
/* Classifies the volume and writes every triangle into a device buffer the
 * session owns; returns the triangle count (synchronises). */
int kf_mesh_generate(kf_session *s, float iso, uint64_t *triangles);
/* Copies min(triangles, max_triangles) triangles, 12 floats each. */
int kf_mesh_read(kf_session *s, float *vertices, uint64_t max_triangles);
/* No counterpart in the reference: kf_mesh_generate for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same triangles, in the same order
 * and with the same bits, written to the caller's device buffer vertices_dev
 * (room for max_triangles triangles of 3 float4 (x, y, z, 0) each, 16-byte
 * aligned); the total triangle count goes to the caller's device word
 * *triangles_dev (8-byte aligned) -- when it exceeds max_triangles only the
 * first max_triangles triangles are written.  Allocates nothing and does not
 * synchronise: it only queues kernels on the session stream (synchronise,
 * e.g. kf_synchronize, before reading the results).  The per-brick counts and
 * offsets live in a fixed device scratch array of about 16 MiB per GPU, set
 * up by kf_session_create; it holds volumes up to 1024 voxels per side
 * (2^21 Morton brick slots) and larger volumes are rejected with
 * KF_ERR_INVALID.  Every call on a GPU shares that GPU's scratch: calls that
 * overlap on one GPU (other streams, other threads) race; calls queued on
 * one stream are ordered.  The session's own mesh (kf_mesh_read /
 * kf_mesh_export_*) is untouched. */
int kf_mesh_generate_noalloc(kf_session *s, float iso, float *vertices_dev, uint64_t max_triangles,
                             uint64_t *triangles_dev);
/* exportMesh (mcubes.cpp:429-509): binary STL of the last kf_mesh_generate,
 * with the reference's layout (zero 80-byte header, a count field of four
 * times the triangle count, normals (1, 1, 1), zero attribute words). */
int kf_mesh_export_stl(kf_session *s, const char *path);
/* exportPointCloud (mcubes.cpp:390-427): ASCII PLY of every vertex. */
int kf_mesh_export_ply(kf_session *s, const char *path);
/* The triangle table: tri[c * 16 + 3 k + v] edge of vertex v of triangle k of
 * cube case c (255 after the last), ntri[c] triangles of case c.  Edge e joins
 * corners (e & 3, (e + 1) & 3) on the bottom face (e < 4), the same on the top
 * face (4 <= e < 8, corners + 4), corner e - 8 to e - 4 otherwise; corner k is
 * at (x, y, z) = {(0,0,0),(1,0,0),(1,0,1),(0,0,1),(0,1,0),(1,1,0),(1,1,1),(0,1,1)}[k]. */
int kf_mc_tables(uint8_t *tri, uint8_t *ntri);

/* ---- multi-session batches -------------------------------------------------
 * No counterpart in the reference (one session per process there).  A batch
 * runs the frame block of many sessions with one launch per kernel type:
 * sessions share depth size, ICP schedule, storage type and options; volume
 * sizes may differ.  Each session's results are the same bits as running it
 * alone with kf_frame_u16_dev.  Creating or re-filling a batch does not block
 * the host: the batch stream waits for work already queued on the members'
 * own streams, and the new tables reach the device in stream order, so frames
 * queued before a membership change still run with the old members.  Call
 * kf_batch_synchronize before using a member session alone again;
 * kf_session_destroy waits for the whole device, so a session may be
 * destroyed once it is no longer a member. */
typedef struct kf_batch kf_batch;
/* Batched sessions track point-to-plane (KF_OPT_TRACK_MODE = 1) and a batch
 * frame does not run recoverPose (as KF_OPT_RECOVER = 0). */
int kf_batch_create(kf_session *const *sessions, int n, void *stream, kf_batch **out); /* n <= 64 */
int kf_batch_set_sessions(kf_batch *b, kf_session *const *sessions, int n);
int kf_batch_destroy(kf_batch *b);
int kf_batch_synchronize(kf_batch *b);
int kf_batch_get_stream(kf_batch *b, void **stream);
/* One frame for every session i with active[i] != 0 (active NULL: all), its
 * 512x424 uint16 mm depth at depth_mm_dev[i] (device memory).  The tick's
 * record (depth pointers, active mask) goes to the device through a ring of
 * four pinned slots, so the call does not block unless four ticks are still
 * queued. */
int kf_batch_frame_u16_dev(kf_batch *b, const uint16_t *const *depth_mm_dev, const int *active);
/* Batch options (no counterpart in the reference; results are the same bits
 * for every setting):
 *   KF_OPT_GRAPH          1: each tick replays a captured hipGraph of its
 *                         record copy and launches (one per ring slot,
 *                         captured again after a membership change); 0
 *                         (default): direct launches, which the host queues
 *                         ahead of the GPU anyway (measured as fast).
 *   KF_OPT_ICP_PERSISTENT 1 (default): Track() of all members in one
 *                         persistent launch, each session on its own group of
 *                         blocks (at most one block per compute unit in all);
 *                         0: one launch per ICP iteration.  Off when a member
 *                         has per-pixel recording on. */
int kf_batch_set_option(kf_batch *b, int option, int value);
int kf_batch_get_option(kf_batch *b, int option, int *value);

const char *kf_last_error(void);
const char *kf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* KFUSION_H */
