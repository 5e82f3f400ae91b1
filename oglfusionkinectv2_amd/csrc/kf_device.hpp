// kf_device.hpp -- device-side math and state for the KinectFusion kernels.
//
// Float expression order follows the reference shaders and glm exactly and the
// library is compiled with -ffp-contract=off, so the FP32 results are the same
// bits as a CPU evaluation of the same expressions:
//   dot(a,b)          = (a.x*b.x + a.y*b.y) + a.z*b.z      (GLSL dot)
//   rotate(M,V)       shaders/integrate.cs:21-27
//   opMul(M,v)        shaders/integrate.cs:29-35
//   M*v (mat4*vec4)   column combination summed left to right
//   A*B (glm mat4)    result[j] = A[0]*B[j][0] + A[1]*B[j][1] + A[2]*B[j][2] + A[3]*B[j][3]
// Matrices are 16 floats, column-major (m[c*4 + r] == glm m[c][r]).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kf {

// Per-session frame state kept in HBM so that the ICP loop, the accept /
// reject decision and the integrate gate never need a host round trip.
struct DevState {
    float pose[16];       // gFusion::m_pose (camera -> world)
    float old_pose[16];   // pose at the start of Track() (gFusion.cpp:653)
    float partials[256];  // 8x32 partial sums of the last iteration (reduce.cs layout)
    double x[6];          // last solve
    float ae;             // alignment energy of the last iteration
    int level_done;       // early exit of the current level (gFusion.cpp:707-708)
    int tracked;          // Track() result
    int integrate_now;    // frame-loop gate (kinect.cpp:677)
    int counter;          // frame counter since reset (kinect.cpp:691)
    int reset;            // reset window flag (kinect.cpp:472, 680-681)
    int iters_run;        // ICP iterations executed in the last Track()
    unsigned int ticket;  // last-block ticket of the fused track+reduce kernel
    // point-to-volume tracking (gFusion.h:426, :445) and the pose library (:424)
    float pose_eig[16];       // m_pose_eig, column-major
    double cum_twist[6];      // m_cumTwist
    float sdf_partials[256];  // 8x32 partial sums of the last TrackSDF iteration (reduceSDF.cs layout)
    float sdf_ae;             // its alignment energy
    int sdf_tracked;          // result of the last TrackSDF() / recoverPose()
    int sdf_iters;            // TrackSDF iterations executed by the last launch
    int lib_n;                // pose library entries
    int lib_overflow;         // a pose was not added because the library was full
    int sdf_error;            // a grid barrier of the persistent TrackSDF kernel timed out
    unsigned int sdf_bar;     // grid barrier arrivals of the persistent TrackSDF kernel
    unsigned int sdf_exit;    // its exit ticket (the last block clears the barrier)
    int sdf_abort;            // set by a block whose barrier wait timed out
    double sdf_log[64][10];   // per iteration: level, iteration, AE, |step|, result[6]
    int icp_error;            // a grid barrier of the persistent Track() kernel timed out
    unsigned int icp_bar;     // its grid barrier arrivals
    unsigned int icp_exit;    // its exit ticket
    int icp_abort;            // set by a block whose barrier wait timed out
    unsigned int icp_epoch;   // launches of that kernel so far (tags its partial rows)
    double pose_d[16];        // FP64 pose of KF_OPT_POSE_FP64 (valid while it casts to `pose`)
};

// Grid barrier of a persistent launch whose blocks are all resident at once:
// arrivals are counted in *bar (this barrier is complete at `target`), the
// wait is bounded by `spin_limit` polls, after which the launch's abort flag
// is raised; a raised flag releases every waiting block.  Returns false when
// the launch is being aborted.  All threads of the block call it.
__device__ __forceinline__ bool grid_barrier(unsigned int *bar, int *abort_flag, unsigned int target,
                                             unsigned int spin_limit, int *sh_abort)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // release the block's stores
        atomicAdd(bar, 1u);
        unsigned int n = 0;
        int ab = 0;
        for (;;) {
            if (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
            if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ab = 1;
                break;
            }
            if (++n > spin_limit) {
                __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ab = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *sh_abort = ab;
    }
    __syncthreads();
    __threadfence();  // acquire the other blocks' stores
    return *sh_abort == 0;
}

// The same barrier for launches whose blocks exchange data only through
// device-coherent accesses (agent-scope atomic stores and loads, which go
// past the per-XCD L2s): each wavefront waits until its own stores are
// complete, then the block arrives.  No L2 write-back or invalidate, so data
// that the blocks only read (written by earlier launches) stays cached
// across barriers.  Ordinary stores made before the barrier are not
// published by it.
__device__ __forceinline__ bool grid_barrier_coherent(unsigned int *bar, int *abort_flag, unsigned int target,
                                                      unsigned int spin_limit, int *sh_abort)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned int n = 0;
        int ab = 0;
        for (;;) {
            if (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
            if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ab = 1;
                break;
            }
            if (++n > spin_limit) {
                __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ab = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *sh_abort = ab;
    }
    __syncthreads();
    return *sh_abort == 0;
}

// Exit ticket of such a launch: the last block out clears the barrier state
// for the next launch.  All threads of the block call it.
__device__ __forceinline__ void grid_exit(unsigned int *bar, unsigned int *exit_count, int *abort_flag,
                                          unsigned int nblk, unsigned int *epoch = nullptr)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned int t = atomicAdd(exit_count, 1u);
        if (t == nblk - 1u) {
            *bar = 0u;
            *abort_flag = 0;
            if (epoch) *epoch += 1u;
            __threadfence();
            *exit_count = 0u;
        }
    }
}

struct Mat16 {
    float m[16];
};

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return (ax * bx + ay * by) + az * bz;
}

__device__ __forceinline__ void rotate3(const float *M, float vx, float vy, float vz, float out[3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = dot3(M[i], M[4 + i], M[8 + i], vx, vy, vz);
}

__device__ __forceinline__ void opmul3(const float *M, float vx, float vy, float vz, float out[3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = dot3(M[i], M[4 + i], M[8 + i], vx, vy, vz) + M[12 + i];
}

__device__ __forceinline__ void matvec4(const float *M, const float v[4], float out[4])
{
#pragma unroll
    for (int i = 0; i < 4; ++i)
        out[i] = ((M[i] * v[0] + M[4 + i] * v[1]) + M[8 + i] * v[2]) + M[12 + i] * v[3];
}

// GLSL min / max semantics (min(x,y) = y < x ? y : x)
__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }

// getCameraMatrix: include/gFusion.h:461-470
__device__ __forceinline__ void camera_matrix(const float cam[4], float K[16])
{
#pragma unroll
    for (int i = 0; i < 16; ++i) K[i] = 0.0f;
    K[0] = cam[0];
    K[5] = cam[1];
    K[8] = cam[2];
    K[9] = cam[3];
    K[10] = 1.0f;
    K[15] = 1.0f;
}

// getInverseCameraMatrix: include/gFusion.h:451-459
__device__ __forceinline__ void inverse_camera_matrix(const float cam[4], float iK[16])
{
#pragma unroll
    for (int i = 0; i < 16; ++i) iK[i] = 0.0f;
    iK[0] = 1.0f / cam[0];
    iK[5] = 1.0f / cam[1];
    iK[8] = -cam[2] / cam[0];
    iK[9] = -cam[3] / cam[1];
    iK[10] = 1.0f;
    iK[15] = 1.0f;
}

// glm mat4 product
__device__ __forceinline__ void mul4(const float *A, const float *B, float *out)
{
    float t[16];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            t[j * 4 + i] = ((A[i] * B[j * 4 + 0] + A[4 + i] * B[j * 4 + 1]) + A[8 + i] * B[j * 4 + 2]) +
                           A[12 + i] * B[j * 4 + 3];
#pragma unroll
    for (int i = 0; i < 16; ++i) out[i] = t[i];
}

// General 4x4 inverse by 2x2 sub-determinants (stands in for glm::inverse,
// gFusion.cpp:747 and :1242).  a(r,c) = M[c*4+r].
__device__ __forceinline__ void inverse4(const float *M, float *out)
{
#define A_(r, c) M[(c) * 4 + (r)]
    const float s0 = A_(0, 0) * A_(1, 1) - A_(1, 0) * A_(0, 1);
    const float s1 = A_(0, 0) * A_(1, 2) - A_(1, 0) * A_(0, 2);
    const float s2 = A_(0, 0) * A_(1, 3) - A_(1, 0) * A_(0, 3);
    const float s3 = A_(0, 1) * A_(1, 2) - A_(1, 1) * A_(0, 2);
    const float s4 = A_(0, 1) * A_(1, 3) - A_(1, 1) * A_(0, 3);
    const float s5 = A_(0, 2) * A_(1, 3) - A_(1, 2) * A_(0, 3);
    const float c5 = A_(2, 2) * A_(3, 3) - A_(3, 2) * A_(2, 3);
    const float c4 = A_(2, 1) * A_(3, 3) - A_(3, 1) * A_(2, 3);
    const float c3 = A_(2, 1) * A_(3, 2) - A_(3, 1) * A_(2, 2);
    const float c2 = A_(2, 0) * A_(3, 3) - A_(3, 0) * A_(2, 3);
    const float c1 = A_(2, 0) * A_(3, 2) - A_(3, 0) * A_(2, 2);
    const float c0 = A_(2, 0) * A_(3, 1) - A_(3, 0) * A_(2, 1);
    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    const float id = 1.0f / det;
    float B[16];
#define B_(r, c, v) B[(c) * 4 + (r)] = (v) * id
    B_(0, 0, A_(1, 1) * c5 - A_(1, 2) * c4 + A_(1, 3) * c3);
    B_(0, 1, -A_(0, 1) * c5 + A_(0, 2) * c4 - A_(0, 3) * c3);
    B_(0, 2, A_(3, 1) * s5 - A_(3, 2) * s4 + A_(3, 3) * s3);
    B_(0, 3, -A_(2, 1) * s5 + A_(2, 2) * s4 - A_(2, 3) * s3);
    B_(1, 0, -A_(1, 0) * c5 + A_(1, 2) * c2 - A_(1, 3) * c1);
    B_(1, 1, A_(0, 0) * c5 - A_(0, 2) * c2 + A_(0, 3) * c1);
    B_(1, 2, -A_(3, 0) * s5 + A_(3, 2) * s2 - A_(3, 3) * s1);
    B_(1, 3, A_(2, 0) * s5 - A_(2, 2) * s2 + A_(2, 3) * s1);
    B_(2, 0, A_(1, 0) * c4 - A_(1, 1) * c2 + A_(1, 3) * c0);
    B_(2, 1, -A_(0, 0) * c4 + A_(0, 1) * c2 - A_(0, 3) * c0);
    B_(2, 2, A_(3, 0) * s4 - A_(3, 1) * s2 + A_(3, 3) * s0);
    B_(2, 3, -A_(2, 0) * s4 + A_(2, 1) * s2 - A_(2, 3) * s0);
    B_(3, 0, -A_(1, 0) * c3 + A_(1, 1) * c1 - A_(1, 2) * c0);
    B_(3, 1, A_(0, 0) * c3 - A_(0, 1) * c1 + A_(0, 2) * c0);
    B_(3, 2, -A_(3, 0) * s3 + A_(3, 1) * s1 - A_(3, 2) * s0);
    B_(3, 3, A_(2, 0) * s3 - A_(2, 1) * s1 + A_(2, 2) * s0);
#undef B_
#undef A_
#pragma unroll
    for (int i = 0; i < 16; ++i) out[i] = B[i];
}

// Wavefront sum over all 64 lanes with DPP moves (CDNA / GFX9 encodings):
// quad butterflies, row rotations, then row_bcast:15 / row_bcast:31 carry the
// row totals up to lane 63, which is broadcast.  Every lane must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xF, false));
}

// the wavefront's sum of v, complete in lane 63
__device__ __forceinline__ float wave_sum_lane63(float v)
{
    v += dpp_move<0xB1, 0xF>(v);  // quad_perm(1,0,3,2)
    v += dpp_move<0x4E, 0xF>(v);  // quad_perm(2,3,0,1)
    v += dpp_move<0x124, 0xF>(v); // row_ror:4
    v += dpp_move<0x128, 0xF>(v); // row_ror:8
    v += dpp_move<0x142, 0xA>(v); // row_bcast:15 into rows 1 and 3
    v += dpp_move<0x143, 0xC>(v); // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ float wave_sum(float v)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wave_sum_lane63(v)), 63));
}

// out[k] = wave_sum(v[k]) for every k, written by lane 63: the same
// additions, with no branch between the terms so that their independent
// DPP chains interleave
template <int N>
__device__ __forceinline__ void wave_sums_store(const float (&v)[N], float *out)
{
    float s[N];
#pragma unroll
    for (int k = 0; k < N; ++k) s[k] = wave_sum_lane63(v[k]);
    if ((threadIdx.x & 63) == 63) {
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = s[k];
    }
}

}  // namespace kf
