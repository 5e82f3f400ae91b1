// kf_kernels.hip -- HIP kernels of the KinectFusion frame loop for gfx950.
//
//   k_pyramid        depth (u16 / f32 mm) -> 3-level depth mip, vertex and
//                    normal maps in one launch      (gFusion.cpp:496-547,
//                    depthToVert.cs:33-72, vertToNorm.cs:9-57)
//   k_icp_iteration  fused track + reduce + combine + FP64 solve + pose update
//                    (gFusion.cpp:669-708, track.cs:47-163, reduce.cs:27-127,
//                    gFusion.cpp:1127-1168); the last launch of Track() also
//                    does the accept / reject and frame-loop gate
//                    (gFusion.cpp:715-723, kinect.cpp:677-691)
//   k_icp_finish     that accept / reject alone (kf_icp_finish)
//   k_integrate_fast / k_integrate  z-column TSDF update (gFusion.cpp:1234-1287,
//                    integrate.cs:52-169) + brick occupancy / zero flags
//   k_brick_dilate   dilated brick flags + 32^3 coarse cells for the raycast
//   k_raycast        ray march + gradient normals (gFusion.cpp:1289-1366, raycast.cs)
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_ext.h>
#include <float.h>
#include <stdint.h>

#include <type_traits>

#include "kf_device.hpp"
#include "kf_dilate.hpp"
#include "kf_internal.hpp"
#include "kf_march.hpp"
#include "kf_store.hpp"

namespace kf {

// ---------------------------------------------------------------------------
// Pyramid
// ---------------------------------------------------------------------------
template <typename DT>
__device__ __forceinline__ float depth0_at(const DT *in, int w, int x, int y)
{
    return (float)in[y * w + x];
}

// depth at level l via the 2x2 box mean of glGenerateMipmap (gFusion.cpp:508)
template <typename DT>
__device__ __forceinline__ float depth_at(const DT *in, int w0, int l, int x, int y)
{
    if (l == 0) return depth0_at(in, w0, x, y);
    if (l == 1) {
        const int X = 2 * x, Y = 2 * y;
        return ((depth0_at(in, w0, X, Y) + depth0_at(in, w0, X + 1, Y)) +
                (depth0_at(in, w0, X, Y + 1) + depth0_at(in, w0, X + 1, Y + 1))) * 0.25f;
    }
    float q[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int X = 4 * x + 2 * (k & 1), Y = 4 * y + 2 * (k >> 1);
        q[k] = ((depth0_at(in, w0, X, Y) + depth0_at(in, w0, X + 1, Y)) +
                (depth0_at(in, w0, X, Y + 1) + depth0_at(in, w0, X + 1, Y + 1))) * 0.25f;
    }
    return ((q[0] + q[1]) + (q[2] + q[3])) * 0.25f;
}

// depthToVert.cs:56 -- (d / 1000) * rotate(invK_l, (x, y, 1))
__device__ __forceinline__ void vertex_of(float d, const float iK[16], int x, int y, float v[3])
{
    float r[3];
    rotate3(iK, (float)x, (float)y, 1.0f, r);
    const float s = d / 1000.0f;
    v[0] = s * r[0];
    v[1] = s * r[1];
    v[2] = s * r[2];
}

__device__ __forceinline__ float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Level 0 runs one wavefront per 8x8 pixel tile and also records the tile's
// (min, max) nonzero depth in metres, which lets integrate pass over z-column
// chunks that no depth can reach; levels 1-2 run one thread per pixel.
template <typename DT>
__device__ __forceinline__ void pyramid_body(const PyramidArgs &a, const DT *__restrict__ in, int b)
{
    // grid order: level 2, level 1, level 0 (block_start[2] = 0)
    int l = 2;
    if (b >= a.block_start[0]) l = 0;
    else if (b >= a.block_start[1]) l = 1;
    const int w = a.w >> l, h = a.h >> l;
    int x, y;
    float dc;
    if (l == 0) {
        const int tile = (b - a.block_start[0]) * 4 + (int)(threadIdx.x >> 6);
        if (tile >= a.tiles_x * a.tiles_y) return;  // whole wavefront
        const int lane = threadIdx.x & 63;
        x = (tile % a.tiles_x) * 8 + (lane & 7);
        y = (tile / a.tiles_x) * 8 + (lane >> 3);
        const bool inside = x < w && y < h;
        dc = inside ? depth_at(in, a.w, 0, x, y) : 0.0f;
        const float dm = dc / 1000.0f;  // integrate.cs:74, once per pixel instead of per voxel
        const bool valid = inside && dm != 0.0f && dm == dm;
        const float mn = wave_min(valid ? dm : INFINITY), mx = wave_max(valid ? dm : -INFINITY);
        if (lane == 0) a.dtiles[tile] = make_float2(mn, mx);
        if (!inside) return;
        a.depth0[y * w + x] = dc;
        a.depth_m[y * w + x] = dm;
    } else {
        const int idx = (b - a.block_start[l]) * 256 + threadIdx.x;
        if (idx >= w * h) return;
        x = idx % w;
        y = idx / w;
        dc = depth_at(in, a.w, l, x, y);
    }
    const int idx = y * w + x;

    float cam[4];
    const float sc = (float)(1 << l);
#pragma unroll
    for (int k = 0; k < 4; ++k) cam[k] = a.cam[k] / sc;  // camPamsDepth / float(1 << i)
    float iK[16];
    inverse_camera_matrix(cam, iK);

    float v[3];
    vertex_of(dc, iK, x, y, v);
    a.vert[l][idx] = make_float4(v[0], v[1], v[2], 0.0f);

    // vertToNorm.cs:20-24: left/up wrap to -1 at the border (out of bounds -> 0),
    // down reads (x-1, y+1)
    float L[3] = {0, 0, 0}, Rr[3], U[3] = {0, 0, 0}, D[3] = {0, 0, 0};
    if (x > 0) vertex_of(depth_at(in, a.w, l, x - 1, y), iK, x - 1, y, L);
    {
        const int rx = min(x + 1, w - 1);
        vertex_of(depth_at(in, a.w, l, rx, y), iK, rx, y, Rr);
    }
    if (y > 0) vertex_of(depth_at(in, a.w, l, x, y - 1), iK, x, y - 1, U);
    if (x > 0) {
        const int dy = min(y + 1, h - 1);
        vertex_of(depth_at(in, a.w, l, x - 1, dy), iK, x - 1, dy, D);
    }
    float4 n;
    if (L[2] == 0.0f || Rr[2] == 0.0f || U[2] == 0.0f || D[2] == 0.0f) {
        n = make_float4(2.0f, 0.0f, 0.0f, 0.0f);
    } else {
        const float dxv[3] = {Rr[0] - L[0], Rr[1] - L[1], Rr[2] - L[2]};
        const float dyv[3] = {D[0] - U[0], D[1] - U[1], D[2] - U[2]};
        const float c0 = dyv[1] * dxv[2] - dyv[2] * dxv[1];
        const float c1 = dyv[2] * dxv[0] - dyv[0] * dxv[2];
        const float c2 = dyv[0] * dxv[1] - dyv[1] * dxv[0];
        const float len = sqrtf(dot3(c0, c1, c2, c0, c1, c2));
        n = make_float4(c0 / len, c1 / len, c2 / len, 1.0f);
    }
    a.norm[l][idx] = n;
}

constexpr int kSelPerThread = 16;  // tiles one thread of the selecting block holds: up to 4096 tiles

__device__ __forceinline__ unsigned block_sum_u(unsigned v, unsigned *sh)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const unsigned r = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return r;
}

// The hybrid raycast's next heavy tiles, by block 0 (256 threads) of a
// single-session pyramid launch, which runs after the session's earlier
// raycasts and before its next one, beside the pyramid's blocks: at most
// heavy_max tiles at or above max(0.7 x the largest count, kRayHeavyFloor),
// the heaviest ones when more qualify; the counts are cleared.  The flags
// and the list are derived from the same register copy of the counts, so
// every flagged tile is in the list (the threshold admits at most heavy_max
// tiles) and each tile is marched exactly once by the next raycast,
// whatever counts were read; the counts themselves only steer the schedule.
__device__ __forceinline__ void select_heavy_tiles(const RaySched &a)
{
    __shared__ unsigned sh[4];
    __shared__ int sh_pos[4];
    // every count read once (all loads issued before use); tiles beyond
    // 256 x kSelPerThread are never marched four lanes per ray
    const int ntiles = a.tiles;
    const int nt = min(ntiles, 256 * kSelPerThread);
    unsigned c[kSelPerThread];
#pragma unroll
    for (int k = 0; k < kSelPerThread; ++k) {
        const int i = (int)threadIdx.x + 256 * k;
        c[k] = i < nt ? a.tile_cost[i] : 0u;
    }
    unsigned m = 0;
#pragma unroll
    for (int k = 0; k < kSelPerThread; ++k) m = max(m, c[k]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    const unsigned mx = max(max(sh[0], sh[1]), max(sh[2], sh[3]));
    __syncthreads();
    // the smallest threshold in [max(0.7 x the largest count, floor),
    // largest + 1] with at most heavy_max tiles at or above it (a bisection:
    // the number at or above a threshold falls as the threshold rises)
    unsigned lo = max((unsigned)kRayHeavyFloor, (mx * 7u + 9u) / 10u), hi = mx + 1u;
    if (lo > hi) lo = hi;
    while (lo < hi) {
        const unsigned mid = (lo + hi) / 2u;
        unsigned n = 0;
#pragma unroll
        for (int k = 0; k < kSelPerThread; ++k) n += c[k] >= mid ? 1u : 0u;
        if (block_sum_u(n, sh) <= (unsigned)a.heavy_max)
            hi = mid;
        else
            lo = mid + 1u;
    }
    const unsigned thr = lo;
    // flags and the list in tile order (a block-wide scan per 256 tiles)
    int base = 0;
#pragma unroll
    for (int k = 0; k < kSelPerThread; ++k) {
        if (256 * k >= nt) break;  // uniform
        const int i = (int)threadIdx.x + 256 * k;
        const bool heavy = i < nt && c[k] >= thr;
        const unsigned long long bm = __ballot(heavy);
        const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
        if ((threadIdx.x & 63) == 0) sh_pos[threadIdx.x >> 6] = __popcll(bm);
        __syncthreads();
        int before = base;
        for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) before += sh_pos[w];
        if (i < nt) {
            a.heavy_flag[i] = heavy ? 1 : 0;
            if (heavy && before + rank < a.heavy_max) a.heavy[1 + before + rank] = i;
            a.tile_cost[i] = 0u;
        }
        base += sh_pos[0] + sh_pos[1] + sh_pos[2] + sh_pos[3];
        __syncthreads();
    }
    for (int i = nt + (int)threadIdx.x; i < ntiles; i += 256) {
        a.heavy_flag[i] = 0;
        a.tile_cost[i] = 0u;
    }
    if (threadIdx.x == 0) a.heavy[0] = min(base, a.heavy_max);
}

template <typename DT>
__global__ __launch_bounds__(256) void k_pyramid(PyramidArgs a, const DT *__restrict__ in)
{
    if (a.sched.tiles > 0 && blockIdx.x == 0) {
        select_heavy_tiles(a.sched);
        return;
    }
    pyramid_body(a, in, blockIdx.x);
}

// ---------------------------------------------------------------------------
// ICP: fused track + reduce + solve + update
// ---------------------------------------------------------------------------

// LDL^T solve of the symmetric 6x6 system in FP64, fully unrolled so that it
// stays in registers.  Returns false when a pivot is not clearly positive
// (singular or indefinite system), leaving the general solve to solve6.
__device__ __forceinline__ bool ldlt6(const double (&A)[36], const double (&b)[6], double (&x)[6])
{
    double L[6][6], D[6], y[6];
    double dmax = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) dmax = fmax(dmax, fabs(A[i * 6 + i]));
    if (!(dmax > 0.0)) return false;
    const double tol = dmax * 1e-12;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        double d = A[j * 6 + j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k] * D[k];
        ok = ok && (d > tol);
        D[j] = d;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double s = A[i * 6 + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k] * D[k];
            L[i][j] = s / d;
        }
    }
    if (!ok) return false;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
        y[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) y[i] /= D[i];
#pragma unroll
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < 6; ++k) s -= L[k][i] * x[k];
        x[i] = s;
    }
    return true;
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// LDS scratch of the one-thread 6x6 solves below (doubles), so that a solve
// does not inflate the register budget of the kernel that calls it.
constexpr int kSolveWs = 120;
// Minimum-norm solve of the symmetric system held in ws[0..35] (A, already
// symmetrised) by cyclic Jacobi eigen-decomposition, singular values below
// 6*eps*sigma_max treated as zero (Eigen's JacobiSVD threshold); V in
// ws[72..107].  One thread.
__device__ void jacobi6(const double *b, double *x, double *ws)
{
    double(*A)[6] = (double(*)[6])ws;
    double(*V)[6] = (double(*)[6])(ws + 72);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < 6; ++i) {
            diag += A[i][i] * A[i][i];
            for (int j = i + 1; j < 6; ++j) off += A[i][j] * A[i][j];
        }
        if (off == 0.0 || off <= 1e-34 * diag) break;
        for (int p = 0; p < 5; ++p)
            for (int q = p + 1; q < 6; ++q) {
                const double apq = A[p][q];
                if (apq == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 6; ++k) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 6; ++k) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 6; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double smax = 0.0;
    for (int i = 0; i < 6; ++i) smax = fmax(smax, fabs(A[i][i]));
    const double thr = fmax(smax * 6.0 * DBL_EPSILON, DBL_MIN);
    for (int k = 0; k < 6; ++k) x[k] = 0.0;
    for (int i = 0; i < 6; ++i) {
        const double lam = A[i][i];
        if (fabs(lam) < thr) continue;
        double proj = 0.0;
        for (int k = 0; k < 6; ++k) proj += V[k][i] * b[k];
        proj /= lam;
        for (int k = 0; k < 6; ++k) x[k] += V[k][i] * proj;
    }
}

// The same contract solved by Jacobi alone (no LDL^T shortcut), for the
// reference-order Track(): A = (C + C^T) / 2 then jacobi6, the operations of
// the CPU restatement of JacobiSVD::solve in the same order.
__device__ void solve6_jacobi(const double *Cin, const double *b, double *x, double *ws)
{
    double(*A)[6] = (double(*)[6])ws;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) A[i][j] = 0.5 * (Cin[i * 6 + j] + Cin[j * 6 + i]);
    jacobi6(b, x, ws);
}

// Minimum-norm least-squares solve of the symmetric 6x6 system (the contract
// of Eigen::JacobiSVD::solve at gFusion.cpp:696-697): LDL^T in FP64 for the
// positive-definite case, cyclic Jacobi eigen-decomposition otherwise.
// Runs on one thread; `ws` holds kSolveWs doubles.
__device__ void solve6(const double *Cin, const double *b, double *x, double *ws)
{
    double(*A)[6] = (double(*)[6])ws;
    double(*L)[6] = (double(*)[6])(ws + 36);
    double *Dg = ws + 108, *y = ws + 114;
    double dmax = 0.0;
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j < 6; ++j) A[i][j] = 0.5 * (Cin[i * 6 + j] + Cin[j * 6 + i]);
        dmax = fmax(dmax, fabs(A[i][i]));
    }
    // LDL^T
    bool spd = dmax > 0.0;
    const double tol = dmax * 1e-12;
    for (int j = 0; j < 6 && spd; ++j) {
        double d = A[j][j];
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k] * Dg[k];
        if (!(d > tol)) {
            spd = false;
            break;
        }
        Dg[j] = d;
        L[j][j] = 1.0;
        for (int i = j + 1; i < 6; ++i) {
            double s = A[i][j];
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k] * Dg[k];
            L[i][j] = s / d;
        }
    }
    if (spd) {
        for (int i = 0; i < 6; ++i) {
            double s = b[i];
            for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
            y[i] = s;
        }
        for (int i = 0; i < 6; ++i) y[i] /= Dg[i];
        for (int i = 5; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < 6; ++k) s -= L[k][i] * x[k];
            x[i] = s;
        }
        return;
    }
    jacobi6(b, x, ws);
}

// cos and sin of an ICP update angle in FP64.  fast and |t| <= 0.5 (every
// ICP increment of a tracked frame): the Taylor series of sin to degree 15
// and of cos to degree 16 in Horner form (truncation error below 1e-19 on
// that interval, rounding within an ulp or two of the library's values);
// otherwise the library routines.
__device__ __forceinline__ void icp_sincos(double t, bool fast, double &c, double &s)
{
    if (fast && fabs(t) <= 0.5) {
        const double t2 = t * t;
        double ps = -1.0 / 1307674368000.0;
        ps = fma(ps, t2, 1.0 / 6227020800.0);
        ps = fma(ps, t2, -1.0 / 39916800.0);
        ps = fma(ps, t2, 1.0 / 362880.0);
        ps = fma(ps, t2, -1.0 / 5040.0);
        ps = fma(ps, t2, 1.0 / 120.0);
        ps = fma(ps, t2, -1.0 / 6.0);
        s = fma(ps * t2, t, t);
        double pc = 1.0 / 20922789888000.0;
        pc = fma(pc, t2, -1.0 / 87178291200.0);
        pc = fma(pc, t2, 1.0 / 479001600.0);
        pc = fma(pc, t2, -1.0 / 3628800.0);
        pc = fma(pc, t2, 1.0 / 40320.0);
        pc = fma(pc, t2, -1.0 / 720.0);
        pc = fma(pc, t2, 1.0 / 24.0);
        pc = fma(pc, t2, -0.5);
        c = fma(pc, t2, 1.0);
    } else {
        c = cos(t);
        s = sin(t);
    }
}

// glm::eulerAngleXYZ(x3,x4,x5) in double, cast to float, translation x0..2,
// then m_pose = delta * m_pose in float (gFusion.cpp:700-705); cs[k] / sn[k]
// are cos / sin of x[3 + k] in double
__device__ __forceinline__ void pose_delta(const double x[6], const double cs[3], const double sn[3], float d[16])
{
    const double ca = cs[0], sa = sn[0];
    const double cb = cs[1], sb = sn[1];
    const double cc = cs[2], sc = sn[2];
    const double Rm[3][3] = {
        {cb * cc, -cb * sc, sb},
        {sa * sb * cc + ca * sc, -sa * sb * sc + ca * cc, -sa * cb},
        {-ca * sb * cc + sa * sc, ca * sb * sc + sa * cc, ca * cb},
    };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int r = 0; r < 3; ++r) d[c * 4 + r] = (float)Rm[r][c];
        d[c * 4 + 3] = 0.0f;
    }
    d[12] = (float)x[0];
    d[13] = (float)x[1];
    d[14] = (float)x[2];
    d[15] = 1.0f;
}

__device__ void pose_update_cs(const double x[6], const double cs[3], const double sn[3], float pose[16])
{
    float d[16];
    pose_delta(x, cs, sn, d);
    mul4(d, pose, pose);
}



// The same update composed in FP64 (the precision variant): delta in double
// times the double pose pd, then the float pose is pd rounded.  pd shadows the
// float pose: when the float pose no longer equals pd rounded (it was set by
// setPose, a reset, a rejected frame or TrackSDF), pd starts again from it.
__device__ void pose_update_cs_d(const double x[6], const double cs[3], const double sn[3], float pose[16],
                                 double pd[16])
{
    bool same = true;
    for (int i = 0; i < 16; ++i) same = same && ((float)pd[i] == pose[i]);
    if (!same)
        for (int i = 0; i < 16; ++i) pd[i] = (double)pose[i];
    const double ca = cs[0], sa = sn[0];
    const double cb = cs[1], sb = sn[1];
    const double cc = cs[2], sc = sn[2];
    const double Rm[3][3] = {
        {cb * cc, -cb * sc, sb},
        {sa * sb * cc + ca * sc, -sa * sb * sc + ca * cc, -sa * cb},
        {-ca * sb * cc + sa * sc, ca * sb * sc + sa * cc, ca * cb},
    };
    double d[16];
    for (int c = 0; c < 3; ++c) {
        for (int r = 0; r < 3; ++r) d[c * 4 + r] = Rm[r][c];
        d[c * 4 + 3] = 0.0;
    }
    d[12] = x[0];
    d[13] = x[1];
    d[14] = x[2];
    d[15] = 1.0;
    double t[16];
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            t[j * 4 + i] = ((d[i] * pd[j * 4 + 0] + d[4 + i] * pd[j * 4 + 1]) + d[8 + i] * pd[j * 4 + 2]) +
                           d[12 + i] * pd[j * 4 + 3];
    for (int i = 0; i < 16; ++i) {
        pd[i] = t[i];
        pose[i] = (float)t[i];
    }
}

// track.cs:47-163 for one pixel, split in two so that a thread can issue the
// loads of several pixels before it waits for any: track_project takes the
// pixel's normal / vertex and finds the reference-map texel (code -1 / -2
// decided here, 0 = pending), track_finish takes that texel and returns the
// result code; error / J are valid for 1.
struct TrackPx {
    int code;
    int ri;  // reference texel, -1 when the projection is NaN (image load out of bounds)
    float pv[3];
    float4 n;
};

__device__ __forceinline__ void track_project(const IcpArgs &a, const float *T, const float *view, float4 n,
                                              float4 v, TrackPx &t)
{
    t.n = n;
    t.ri = -1;
    if (n.x == 2.0f) {
        t.code = -1;
        return;
    }
    const float v4[4] = {v.x, v.y, v.z, 1.0f};
    float pv[4], pp[4];
    matvec4(T, v4, pv);
    matvec4(view, pv, pp);
    t.pv[0] = pv[0];
    t.pv[1] = pv[1];
    t.pv[2] = pv[2];
    const float px = pp[0] / pp[2] + 0.5f;
    const float py = pp[1] / pp[2] + 0.5f;
    if (px < 0.0f || px > (float)(a.rw - 1) || py < 0.0f || py > (float)(a.rh - 1)) {
        t.code = -2;
        return;
    }
    // NaN coordinates pass the bounds test; their image load is out of bounds (zero)
    if (px == px && py == py) t.ri = (int)py * a.rw + (int)px;
    t.code = 0;
}

__device__ __forceinline__ int track_finish(const IcpArgs &a, const float *T, const TrackPx &t, float4 N4, float4 V4,
                                            float &err, float J[6])
{
    if (t.code != 0) return t.code;
    float rn[3] = {0.0f, 0.0f, 0.0f}, rv[3] = {0.0f, 0.0f, 0.0f};
    if (t.ri >= 0) {
        rn[0] = N4.x; rn[1] = N4.y; rn[2] = N4.z;
        rv[0] = V4.x; rv[1] = V4.y; rv[2] = V4.z;
    }
    if (rn[0] == -2.0f) return -3;
    const float *pv = t.pv;
    const float diff[3] = {rv[0] - pv[0], rv[1] - pv[1], rv[2] - pv[2]};
    float pn[3];
    rotate3(T, t.n.x, t.n.y, t.n.z, pn);
    if (sqrtf(dot3(diff[0], diff[1], diff[2], diff[0], diff[1], diff[2])) > a.dist_th) return -4;
    if (dot3(pn[0], pn[1], pn[2], rn[0], rn[1], rn[2]) < a.normal_th) return -5;
    err = dot3(rn[0], rn[1], rn[2], diff[0], diff[1], diff[2]);
    J[0] = rn[0];
    J[1] = rn[1];
    J[2] = rn[2];
    J[3] = pv[1] * rn[2] - pv[2] * rn[1];
    J[4] = pv[2] * rn[0] - pv[0] * rn[2];
    J[5] = pv[0] * rn[1] - pv[1] * rn[0];
    return 1;
}

constexpr int kIcpThreads = 256;
#ifndef KF_ICP_PIXELS
#define KF_ICP_PIXELS 4
#endif
constexpr int kIcpPx = KF_ICP_PIXELS;  // pixels whose loads a thread issues together
constexpr int kTerms = 28;  // e^2, e*J[6], upper-triangle J J^T [21]
constexpr int kPartStride = 32;  // floats per block-partial row (16-byte aligned rows)

// ICP iteration pieces (single-session and batched kernels).
//
// Block b serves row group g = b % 8 (the reduce.cs work group that owns rows
// y = g mod 8) and takes every rows_step-th row of that group.  Per-thread FP32
// products, DPP wave sums, block partials in FP32; the partials of each group
// are then summed in FP64, the combine rule applied, the 6x6 solved and the
// pose updated.

// getCameraMatrix(camPams) * inverse(m_pose) (gFusion.cpp:747)
__device__ __forceinline__ void icp_view(const IcpArgs &a, const float T[16], float view[16])
{
    float K[16], ip[16];
    camera_matrix(a.cam, K);
    inverse4(T, ip);
    mul4(K, ip, view);
}

// pixels of this thread in a block's rows (one chunk of kIcpPx at most when
// icp_one_chunk holds)
__device__ __forceinline__ int icp_block_pixels(const IcpArgs &a, int g, int jb)
{
    const int y0 = g + 8 * jb, ystep = 8 * a.rows_step;
    const int per_row = (a.w + kIcpThreads - 1) / kIcpThreads;
    const int nrows = y0 < a.h ? (a.h - 1 - y0) / ystep + 1 : 0;
    return nrows * per_row;
}

// the map texels of the thread's first kIcpPx pixels (normal (2, 0, 0, 0) and
// vertex 0 beyond its pixels): what icp_block_partials loads first
__device__ __forceinline__ void icp_load_pixels(const IcpArgs &a, int g, int jb, float4 (&nv)[kIcpPx],
                                                float4 (&vv)[kIcpPx])
{
    const int y0 = g + 8 * jb, ystep = 8 * a.rows_step;
    const int per_row = (a.w + kIcpThreads - 1) / kIcpThreads;
    const int npx = icp_block_pixels(a, g, jb);
#pragma unroll
    for (int p = 0; p < kIcpPx; ++p) {
        const int r = p / per_row, m = p - r * per_row;
        const int x = threadIdx.x + kIcpThreads * m, y = y0 + ystep * r;
        nv[p] = make_float4(2.0f, 0.0f, 0.0f, 0.0f);
        vv[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (p < npx && x < a.w) {
            nv[p] = a.norm[y * a.w + x];
            vv[p] = a.vert[y * a.w + x];
        }
    }
}

// this block's 28 sums at pose T written to dst[0..27] (threads < 28); ends
// with the block synchronised.  g: row group, jb: block within the group.
// CACHED: the block has at most kIcpPx pixels per thread and cn / cv hold
// their map texels (icp_load_pixels), kept over the iterations of a level.
// dst64: the row is written as tagged words, (tag << 32) | float bits, for
// readers that poll it (icp_group_sums_tagged); dst is then unused.
template <bool CACHED = false>
__device__ __forceinline__ void icp_block_partials(const IcpArgs &a, const float T[16], int g, int jb, float *dst,
                                                   float (*wsum)[kTerms], const float4 *cn = nullptr,
                                                   const float4 *cv = nullptr, unsigned long long *dst64 = nullptr,
                                                   unsigned tag = 0u)
{
    float view[16];
    icp_view(a, T, view);

    float acc[kTerms];
#pragma unroll
    for (int k = 0; k < kTerms; ++k) acc[k] = 0.0f;

    // the thread's pixels in row-major order (rows y = g + 8 jb + 8 rows_step r,
    // columns x = threadIdx.x + kIcpThreads m), kIcpPx at a time: all their
    // map loads, then all their reference gathers, then the sums in order
    const int y0 = g + 8 * jb, ystep = 8 * a.rows_step;
    const int per_row = (a.w + kIcpThreads - 1) / kIcpThreads;
    const int nrows = y0 < a.h ? (a.h - 1 - y0) / ystep + 1 : 0;
    const int npx = nrows * per_row;
    for (int j0 = 0; j0 < npx; j0 += kIcpPx) {
        int idx[kIcpPx];
        TrackPx t[kIcpPx];
#pragma unroll
        for (int p = 0; p < kIcpPx; ++p) {
            const int j = j0 + p, r = j / per_row, m = j - r * per_row;
            const int x = threadIdx.x + kIcpThreads * m, y = y0 + ystep * r;
            idx[p] = (j < npx && x < a.w) ? y * a.w + x : -1;
        }
        float4 nv[kIcpPx], vv[kIcpPx];
#pragma unroll
        for (int p = 0; p < kIcpPx; ++p) {
            nv[p] = make_float4(2.0f, 0.0f, 0.0f, 0.0f);
            vv[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (CACHED) {
                nv[p] = cn[p];
                vv[p] = cv[p];
            } else if (idx[p] >= 0) {
                nv[p] = a.norm[idx[p]];
                vv[p] = a.vert[idx[p]];
            }
        }
#pragma unroll
        for (int p = 0; p < kIcpPx; ++p) track_project(a, T, view, nv[p], vv[p], t[p]);
        float4 N4[kIcpPx], V4[kIcpPx];
#pragma unroll
        for (int p = 0; p < kIcpPx; ++p) {
            N4[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            V4[p] = N4[p];
            if (idx[p] >= 0 && t[p].code == 0 && t[p].ri >= 0) {
                N4[p] = a.refN[t[p].ri];
                V4[p] = a.refV[t[p].ri];
            }
        }
#pragma unroll
        for (int p = 0; p < kIcpPx; ++p) {
            if (idx[p] < 0) continue;
            float e = 0.0f, J[6] = {0, 0, 0, 0, 0, 0};
            const int r = track_finish(a, T, t[p], N4[p], V4[p], e, J);
            if (a.dbg_result) {
                const int i = idx[p];
                a.dbg_result[i] = r;
                if (r == 1) {
                    a.dbg_error[i] = e;
#pragma unroll
                    for (int k = 0; k < 6; ++k) a.dbg_J[6 * i + k] = J[k];
                }
            }
            if (r == 1) {
                acc[0] += e * e;
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[1 + k] += e * J[k];
                int q = 7;
#pragma unroll
                for (int p2 = 0; p2 < 6; ++p2)
#pragma unroll
                    for (int s = p2; s < 6; ++s) acc[q++] += J[p2] * J[s];
            }
        }
    }

    wave_sums_store(acc, wsum[threadIdx.x >> 6]);
    __syncthreads();
    if (threadIdx.x < kTerms) {
        const int k = threadIdx.x;
        const float v = (wsum[0][k] + wsum[1][k]) + (wsum[2][k] + wsum[3][k]);
        // device-coherent store: other blocks read the row after a grid
        // barrier that does not write back or invalidate the L2, or poll it
        if (dst64)
            __hip_atomic_store(&dst64[k], ((unsigned long long)tag << 32) | __float_as_uint(v), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else
            __hip_atomic_store(&dst[k], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
}

// FP64 sums per row group of the partial rows 0..nrows-1 into gsum (row
// 8 jb + g holds block jb of row group g; with group_only >= 0 only that
// group's rows were written and the other groups sum to zero).  For each
// group, thread (j, k) = (tid / 32, tid % 32) sums term k of the rows
// jb = j, j + 8, j + 16, ... in order (coalesced device-coherent loads, all
// issued before the adds), the two halves of a wavefront are added with one
// xor-32 shuffle and the four waves through LDS: a fixed order, the same in
// every block and for a group whether or not the other groups are summed.
// Ends with the block synchronised.
__device__ __forceinline__ void icp_group_sums(const float *partials, int nrows, int group_only,
                                               double (*wgsum)[8][kTerms], double (*gsum)[kTerms])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = threadIdx.x >> 5, k = threadIdx.x & 31;
    const int per_group = nrows >> 3;
    for (int g = 0; g < 8; ++g) {
        if (group_only >= 0 && g != group_only) {  // uniform
            if (lane < kTerms) wgsum[wave][g][lane] = 0.0;
            continue;
        }
        double ps = 0.0;
        if (k < kTerms) {
            for (int jb0 = j; jb0 < per_group; jb0 += 64) {
                float q[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int jb = jb0 + 8 * u;
                    q[u] = jb < per_group ? __hip_atomic_load(partials + (8 * jb + g) * kPartStride + k,
                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                          : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (jb0 + 8 * u < per_group) ps += (double)q[u];
            }
        }
        ps += __shfl_xor(ps, 32, 64);
        if (lane < kTerms) wgsum[wave][g][lane] = ps;
    }
    __syncthreads();
    if (threadIdx.x < 8 * kTerms) {
        const int g = threadIdx.x / kTerms, t = threadIdx.x % kTerms;
        gsum[g][t] = (wgsum[0][g][t] + wgsum[1][g][t]) + (wgsum[2][g][t] + wgsum[3][g][t]);
    }
    __syncthreads();
}

// icp_group_sums over rows of tagged words that other blocks of the same
// launch are writing: a term is taken once its word carries `tag` (polled
// device-coherent loads; no barrier, and no ordering between words is
// needed since each word holds its own tag).  Same order as icp_group_sums,
// so the same bits.  Returns false when a wait exceeded spin_limit polls
// (the launch's abort flag is then raised) or the flag was raised by another
// block.  Ends with the block synchronised.
__device__ __forceinline__ bool icp_group_sums_tagged(unsigned long long *rows, int nrows, int group_only,
                                                      unsigned tag, unsigned spin_limit, int *abort_flag,
                                                      double (*wgsum)[8][kTerms], double (*gsum)[kTerms])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = threadIdx.x >> 5, k = threadIdx.x & 31;
    const int per_group = nrows >> 3;
    bool ab = false;
    for (int g = 0; g < 8; ++g) {
        if (group_only >= 0 && g != group_only) {  // uniform
            if (lane < kTerms) wgsum[wave][g][lane] = 0.0;
            continue;
        }
        double ps = 0.0;
        if (k < kTerms) {
            for (int jb0 = j; jb0 < per_group; jb0 += 64) {
                unsigned long long q[8];
                unsigned pending = 0u;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int jb = jb0 + 8 * u;
                    q[u] = 0ull;
                    if (jb < per_group) {
                        q[u] = __hip_atomic_load(rows + (8 * jb + g) * kPartStride + k, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
                        if ((unsigned)(q[u] >> 32) != tag) pending |= 1u << u;
                    }
                }
                unsigned n = 0;
                while (pending != 0u && !ab) {
                    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        ab = true;
                        break;
                    }
                    if (++n > spin_limit) {
                        __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ab = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (pending & (1u << u)) {
                            q[u] = __hip_atomic_load(rows + (8 * (jb0 + 8 * u) + g) * kPartStride + k,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            if ((unsigned)(q[u] >> 32) == tag) pending &= ~(1u << u);
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (jb0 + 8 * u < per_group) ps += (double)__uint_as_float((unsigned)q[u]);
            }
        }
        ps += __shfl_xor(ps, 32, 64);
        if (lane < kTerms) wgsum[wave][g][lane] = ps;
    }
    const int any_ab = __syncthreads_or(ab ? 1 : 0);
    if (threadIdx.x < 8 * kTerms) {
        const int g = threadIdx.x / kTerms, t = threadIdx.x % kTerms;
        gsum[g][t] = (wgsum[0][g][t] + wgsum[1][g][t]) + (wgsum[2][g][t] + wgsum[3][g][t]);
    }
    __syncthreads();
    return any_ab == 0;
}

// exposed 8x32 partials in the reduce.cs layout (one slot per thread);
// reduce.cs counts every pixel of the group's rows once (:47-62, :101)
__device__ __forceinline__ void icp_expose(const IcpArgs &a, double (*gsum)[kTerms], float *out)
{
    const int g = threadIdx.x >> 5, k = threadIdx.x & 31;
    const int rows = (a.h > g) ? (a.h - g + 7) / 8 : 0;
    float val = 0.0f;
    if (k < kTerms) val = (float)gsum[g][k];
    else if (k == kTerms) val = (float)((double)rows * (double)a.w);
    out[threadIdx.x] = val;
}

// Combine rule, AE (gFusion.cpp:1143-1167) and the 6x6 solve on a whole
// wavefront (every lane returns the same x and AE).  Returns |x|.
__device__ __forceinline__ double icp_solve(const IcpArgs &a, double (*gsum)[kTerms], double *solve_ws,
                                            double (&x)[6], float &ae)
{
    const int lane = threadIdx.x & 63;
    double v[kTerms], count = 0.0;
#pragma unroll
    for (int k = 0; k < kTerms; ++k) v[k] = 0.0;
    if (a.rule == 1) {  // sum all rows
#pragma unroll
        for (int g = 0; g < 8; ++g) {
#pragma unroll
            for (int k = 0; k < kTerms; ++k) v[k] += gsum[g][k];
            count += (double)((a.h > g) ? (a.h - g + 7) / 8 : 0) * (double)a.w;
        }
    } else {  // reference: rows 1..7 assigned over row 0, only row 7 survives
#pragma unroll
        for (int k = 0; k < kTerms; ++k) v[k] = gsum[7][k];
        count = (double)((a.h > 7) ? (a.h - 7 + 7) / 8 : 0) * (double)a.w;
    }
    ae = sqrtf((float)v[0] / (float)count);  // gFusion.cpp:1167
    // every lane solves the whole system from its own copy of the sums (the
    // same bits on every lane; no cross-lane traffic on the solve's chain)
    double A[36], bb[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) bb[k] = v[1 + k];
    {
        int q = 7;
#pragma unroll
        for (int p = 0; p < 6; ++p)
#pragma unroll
            for (int s2 = p; s2 < 6; ++s2) {
                A[p * 6 + s2] = v[q];
                A[s2 * 6 + p] = v[q];
                ++q;
            }
    }
    if (!ldlt6(A, bb, x)) {
        // singular / indefinite: minimum-norm solve on lane 0 in LDS
        double *Cl = solve_ws + kSolveWs, *bl = Cl + 36, *xl = bl + 6;
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < 36; ++k) Cl[k] = A[k];
#pragma unroll
            for (int k = 0; k < 6; ++k) bl[k] = bb[k];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (lane == 0) solve6(Cl, bl, xl, solve_ws);
#pragma unroll
        for (int k = 0; k < 6; ++k) x[k] = readlane_d(lane == 0 ? xl[k] : 0.0, 0);
    }
    double nrm = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) nrm += x[k] * x[k];
    return sqrt(nrm);
}

// Track() accept / reject (gFusion.cpp:715-723); with `frame` the frame-loop
// gate of kinect.cpp:677-691: integrate if tracked or within the reset window.
// One lane.
__device__ __forceinline__ void icp_finish_body(DevState *st, int frame)
{
    const float ae = st->ae;
    int tracked;
    if (ae > 0.5f || ae == 0.0f) {
        for (int i = 0; i < 16; ++i) st->pose[i] = st->old_pose[i];
        tracked = 0;
    } else {
        tracked = 1;
    }
    st->tracked = tracked;
    if (frame) {
        const int go = tracked || st->reset;
        st->integrate_now = go;
        if (go && st->counter > 2) st->reset = 0;
        st->counter += 1;
    }
}

// The accept / reject above on lane 0, then, when tracked, updatePoseFinder
// (gFusion.cpp:722, 1570-1597) on the whole wavefront: the pose joins the
// library unless its translation lies within 0.2 m (glm::l2Norm) of an
// entry's.  All 64 lanes of one wavefront call this.
__device__ __forceinline__ void icp_finish_wave(DevState *st, int frame, float *lib, int lib_max)
{
    const int lane = threadIdx.x & 63;
    int tracked = 0, n = 0;
    float P[16];
    if (lane == 0) {
        icp_finish_body(st, frame);
        tracked = st->tracked;
        n = st->lib_n;
#pragma unroll
        for (int i = 0; i < 16; ++i) P[i] = st->pose[i];
    }
    tracked = __shfl(tracked, 0, 64);
    if (!tracked || !lib) return;
    n = __shfl(n, 0, 64);
    // lane i < 16 keeps pose element i
    float mine = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const float v = __shfl(lane == 0 ? P[i] : 0.0f, 0, 64);
        if (lane == i) mine = v;
    }
    const float px = __shfl(mine, 12, 64), py = __shfl(mine, 13, 64), pz = __shfl(mine, 14, 64);
    bool near = false;
    for (int i = lane; i < n; i += 64) {
        const float *e = lib + 16 * i;
        const float d0 = px - e[12], d1 = py - e[13], d2 = pz - e[14];
        near = near || (sqrtf(dot3(d0, d1, d2, d0, d1, d2)) < 0.2f);
    }
    if (__any(near)) return;
    if (n < lib_max) {
        if (lane < 16) lib[16 * n + lane] = mine;
        if (lane == 0) st->lib_n = n + 1;
    } else if (lane == 0) {
        st->lib_overflow = 1;
    }
}

// The end of one ICP iteration on the first wavefront (every lane holds the
// same x, |x| and AE): pose update (lanes 0-2 evaluate the three angles' cos
// / sin in parallel, the same double routines, so the same bits), the early-
// exit flag of gFusion.cpp:707-708, and on the last launch of Track() the
// accept / reject.
__device__ __forceinline__ void icp_apply(const IcpArgs &a, DevState *__restrict__ st, const double (&x)[6],
                                          double nrm, float ae)
{
    const int lane = threadIdx.x & 63;
    const double my_ang = lane == 1 ? x[4] : lane == 2 ? x[5] : x[3];
    double c1, s1;
    icp_sincos(my_ang, a.fast_trig != 0, c1, s1);
    const double cs[3] = {c1, __shfl(c1, 1, 64), __shfl(c1, 2, 64)};
    const double sn[3] = {s1, __shfl(s1, 1, 64), __shfl(s1, 2, 64)};
    if (lane == 0) {
        float P[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) P[i] = st->pose[i];
        if (a.save_pose) {
#pragma unroll
            for (int i = 0; i < 16; ++i) st->old_pose[i] = P[i];
        }
        if (a.pose_fp64) {
            double Pd[16];
            for (int i = 0; i < 16; ++i) Pd[i] = st->pose_d[i];
            pose_update_cs_d(x, cs, sn, P, Pd);
            for (int i = 0; i < 16; ++i) st->pose_d[i] = Pd[i];
        } else {
            pose_update_cs(x, cs, sn, P);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) st->pose[i] = P[i];
#pragma unroll
        for (int k = 0; k < 6; ++k) st->x[k] = x[k];
        st->ae = ae;
        st->level_done = (nrm < 1e-5 && nrm != 0.0) ? 1 : 0;
        st->iters_run = (a.save_pose ? 0 : st->iters_run) + 1;
        st->ticket = 0u;
    }
    if (a.finish) icp_finish_wave(st, a.frame, a.library, a.lib_max);
}

// One ICP iteration per launch (track.cs + reduce.cs + getReduction + solve +
// pose update); the last block to finish (ticket) does the FP64 part.
// blk / nblk: this block's index and the block count of the session's launch
__device__ __forceinline__ void icp_body(const IcpArgs &a, DevState *__restrict__ st, int blk, int nblk)
{
    __shared__ float wsum[kIcpThreads / 64][kTerms];
    __shared__ int last;
    __shared__ double gsum[8][kTerms];
    __shared__ double wgsum[kIcpThreads / 64][8][kTerms];
    __shared__ double solve_ws[kSolveWs + 36 + 6 + 6];

    if (!a.first && st->level_done) {  // early exit of this level (uniform)
        if (a.finish && blk == 0 && threadIdx.x < 64) icp_finish_wave(st, a.frame, a.library, a.lib_max);
        return;
    }
    float T[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) T[i] = st->pose[i];
    // all groups: block b serves group b % 8; one group g: block b serves g
    // and writes partial row 8 b + g (the row that group's blocks would use)
    const int g = a.group_only >= 0 ? a.group_only : (blk & 7);
    const int jb = a.group_only >= 0 ? blk : (blk >> 3);
    icp_block_partials(a, T, g, jb, a.block_partials + (8 * jb + g) * kPartStride, wsum);
    if (threadIdx.x == 0) {
        // one agent-scope release per block: the barrier above orders the
        // block's partial stores before it
        __threadfence();
        const unsigned t = atomicAdd(&st->ticket, 1u);
        last = (t == (unsigned)(nblk - 1));
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // acquire: every block's partials are visible
    icp_group_sums(a.block_partials, 8 * a.rows_step, a.group_only, wgsum, gsum);
    icp_expose(a, gsum, st->partials);
    if (threadIdx.x >= 64) return;
    double x[6] = {0, 0, 0, 0, 0, 0};
    float ae = 0.0f;
    const double nrm = icp_solve(a, gsum, solve_ws, x, ae);
    icp_apply(a, st, x, nrm, ae);
}

__global__ __launch_bounds__(kIcpThreads) void k_icp_iteration(IcpArgs a, DevState *__restrict__ st)
{
    icp_body(a, st, blockIdx.x, gridDim.x);
}

__global__ void k_icp_finish(DevState *st, int frame, float *lib, int lib_max)
{
    if (threadIdx.x >= 64 || blockIdx.x != 0) return;
    icp_finish_wave(st, frame, lib, lib_max);
}

// ---------------------------------------------------------------------------
// Track() with the reference's summation order (KF_OPT_ICP_REFERENCE_ORDER).
// Per iteration: k_icp_records writes track.cs's per-pixel output
// ({result, error, J[6]}, track.cs:18-27) for every pixel of the level; then
// k_icp_reduce_ref forms reduce.cs's sums in reduce.cs's own order -- block g
// of 8 over rows y = g (mod 8), thread s of 112 over columns x = s (mod 112),
// serial FP32 sums per thread in row-major order, the count term counting
// every pixel, then threads 0..31 adding the 112 per-thread rows serially
// (reduce.cs:27-127) -- and its last block applies getReduction's combine in
// FP32 (gFusion.cpp:1137-1167) and the JacobiSVD-contract solve by Jacobi
// alone.  These are the operations of the CPU restatement in the same order,
// so partials, AE and pose come out as the same bits; the fused kernels above
// sum the same products in another order (wave trees, FP64 group sums).
constexpr int kRefThreads = 112;  // reduce.cs work-group size (gFusion.cpp:795)

__global__ __launch_bounds__(256) void k_icp_records(IcpArgs a, const DevState *__restrict__ st)
{
    if (!a.first && st->level_done) return;  // early exit of this level (uniform)
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.w * a.h) return;
    float T[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) T[k] = st->pose[k];
    float view[16];
    icp_view(a, T, view);
    TrackPx t;
    track_project(a, T, view, a.norm[i], a.vert[i], t);
    float4 N4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), V4 = N4;
    if (t.code == 0 && t.ri >= 0) {
        N4 = a.refN[t.ri];
        V4 = a.refV[t.ri];
    }
    float e = 0.0f, J[6] = {0, 0, 0, 0, 0, 0};
    const int r = track_finish(a, T, t, N4, V4, e, J);
    a.dbg_result[i] = r;
    a.dbg_error[i] = e;
#pragma unroll
    for (int k = 0; k < 6; ++k) a.dbg_J[6 * i + k] = J[k];
}

__global__ __launch_bounds__(128) void k_icp_reduce_ref(IcpArgs a, DevState *__restrict__ st)
{
    __shared__ float S[kRefThreads][32];
    __shared__ int last;
    __shared__ double solve_ws[kSolveWs + 36 + 6 + 6];
    if (!a.first && st->level_done) {  // early exit of this level (uniform)
        if (a.finish && blockIdx.x == 0 && threadIdx.x < 64) icp_finish_wave(st, a.frame, a.library, a.lib_max);
        return;
    }
    const int g = blockIdx.x, sl = threadIdx.x;
    if (sl < kRefThreads) {
        float sums[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) sums[k] = 0.0f;
        for (int y = g; y < a.h; y += 8)
            for (int x = sl; x < a.w; x += kRefThreads) {
                const int i = y * a.w + x;
                const int r = a.dbg_result[i];
                if (r < 1) {  // codes -1 .. -5: each counted once (reduce.cs:47-62)
                    sums[28] += 1.0f;
                    continue;
                }
                const float e = a.dbg_error[i];
                float J[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) J[k] = a.dbg_J[6 * i + k];
                sums[0] += e * e;
#pragma unroll
                for (int k = 0; k < 6; ++k) sums[k + 1] += e * J[k];
                int q = 7;
#pragma unroll
                for (int p = 0; p < 6; ++p)
#pragma unroll
                    for (int s2 = p; s2 < 6; ++s2) sums[q++] += J[p] * J[s2];
                sums[28] += 1.0f;
            }
#pragma unroll
        for (int k = 0; k < 32; ++k) S[sl][k] = sums[k];
    }
    __syncthreads();
    if (sl < 32) {
        float acc = S[0][sl];
        for (int k = 1; k < kRefThreads; ++k) acc += S[k][sl];
        __hip_atomic_store(&st->partials[g * 32 + sl], acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned t = atomicAdd(&st->ticket, 1u);
        last = (t == 7u);
    }
    __syncthreads();
    if (!last || threadIdx.x >= 64) return;
    __threadfence();
    const int lane = threadIdx.x;
    double x[6] = {0, 0, 0, 0, 0, 0};
    float ae = 0.0f;
    double nrm = 0.0;
    if (lane == 0) {
        float P[256];
        for (int k = 0; k < 256; ++k)
            P[k] = __hip_atomic_load(&st->partials[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // getReduction (gFusion.cpp:1137-1167): last-row rule keeps work group
        // 7's row; the sum rule adds the rows in order, in FP32
        float row[32];
        for (int k = 0; k < 32; ++k) {
            if (a.rule == 1) {
                float acc = P[k];
                for (int r = 1; r < 8; ++r) acc += P[r * 32 + k];
                row[k] = acc;
            } else {
                row[k] = P[7 * 32 + k];
            }
        }
        ae = sqrtf(row[0] / row[28]);
        double *C = solve_ws + kSolveWs, *b = C + 36, *xs = b + 6;
        for (int k = 0; k < 6; ++k) b[k] = (double)row[1 + k];
        int q = 7;
        for (int p = 0; p < 6; ++p)
            for (int s2 = p; s2 < 6; ++s2) {
                C[p * 6 + s2] = (double)row[q];
                C[s2 * 6 + p] = (double)row[q];
                ++q;
            }
        solve6_jacobi(C, b, xs, solve_ws);
        for (int k = 0; k < 6; ++k) x[k] = xs[k];
        for (int k = 0; k < 6; ++k) nrm += x[k] * x[k];
        nrm = sqrt(nrm);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = readlane_d(x[k], 0);
    nrm = readlane_d(nrm, 0);
    ae = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ae), 0));
    icp_apply(a, st, x, nrm, ae);
}

// Track() (gFusion.cpp:648-740) in one launch: every level (coarse to fine)
// and iteration.  A block computes the partial row k_icp_iteration's block of
// the same index would (blocks beyond the level's count only read) and
// writes it as tagged words; every block then polls all rows until each word
// carries the iteration's tag (this exchange is the iteration's only
// synchronisation), forms the same FP64 group sums from the same rows in the
// same order, solves the same system and updates the same pose, so the
// blocks agree on the pose and on the early exit (gFusion.cpp:707-708), and
// an exited level costs nothing.  Partial rows alternate between two
// buffers: a block can be at most one iteration ahead of the slowest reader.
// Results are those of the k_icp_iteration sequence bit for bit.
// The launch body of Track() for one session whose grid is `nblk` blocks,
// this one being `blk`: a level of nb blocks of the per-iteration launch
// (virtual blocks) is covered by the session's blocks in turn, block blk
// computing the partial rows of virtual blocks blk, blk + nblk, ...; the rows,
// and so the sums, pose and decisions, are those of the per-iteration
// launches whatever nblk is.  With nblk >= nb a block keeps its one virtual
// block's texels in registers over the level.
__device__ __forceinline__ void icp_track_body(const IcpTrackArgs &a, DevState *__restrict__ st, const int blk,
                                               const int nblk)
{
    __shared__ float wsum[kIcpThreads / 64][kTerms];
    __shared__ double gsum[8][kTerms];
    __shared__ double wgsum[kIcpThreads / 64][8][kTerms];
    __shared__ double solve_ws[kSolveWs + 36 + 6 + 6];
    __shared__ float Ts[16];
    __shared__ double Tsd[16];  // FP64 pose (KF_OPT_POSE_FP64)
    __shared__ int sh_abort, sh_done;
    __shared__ __attribute__((aligned(16))) float solo_rows[8 * kPartStride];

    const int tid = threadIdx.x, lane = tid & 63;
    if (tid < 16) {
        Ts[tid] = st->pose[tid];
        Tsd[tid] = st->pose_d[tid];
    }
    if (tid == 0) {
        sh_abort = 0;
        sh_done = 0;
    }
    __syncthreads();
    if (blk == 0 && tid < 16) st->old_pose[tid] = Ts[tid];  // oldPose (gFusion.cpp:653)

    // partial rows of exchange n carry the tag (launch << 8) + n + 1 (at
    // most 3 x 64 exchanges per launch), so a row is current exactly when its
    // words carry this exchange's tag.  Tags repeat after 2^24 launches: a
    // stale row passes for a current one only if it was written exactly
    // 2^24 k launches earlier at the same exchange and never since (a change
    // of the row layout, kf_set_option, clears the rows).
    const unsigned epoch = __hip_atomic_load(&st->icp_epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned phase = 0, xchg = 0;  // grid barriers and row exchanges so far
    int iters_run = 0;
    bool aborted = false, solo_before = false;
    for (int level = 2; level >= 0 && !aborted; --level) {
        const IcpArgs L = a.lv[level];
        const int nb = icp_blocks(L);
        if (a.iters[level] <= 0) continue;
        // a level served by one block is run by block 0 alone, with no grid
        // barrier (its partial row stays in LDS); the other blocks go on to
        // wait for its pose at the next level that needs them
        const bool solo = nb == 1;
        if (!solo && solo_before) {
            if (blk == 0 && tid < 16) {
                __hip_atomic_store(&st->pose[tid], Ts[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(reinterpret_cast<unsigned long long *>(st->pose_d) + tid,
                                   (unsigned long long)__double_as_longlong(Tsd[tid]), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
            if (!grid_barrier_coherent(&st->icp_bar, &st->icp_abort, (phase + 1u) * (unsigned)nblk, a.spin_limit,
                                       &sh_abort)) {
                aborted = true;
                break;
            }
            ++phase;
            if (tid < 16) {
                Ts[tid] = __hip_atomic_load(&st->pose[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                Tsd[tid] = __longlong_as_double((long long)__hip_atomic_load(
                    reinterpret_cast<unsigned long long *>(st->pose_d) + tid, __ATOMIC_RELAXED,
                    __HIP_MEMORY_SCOPE_AGENT));
            }
            __syncthreads();
        }
        solo_before = solo;
        if (solo && blk != 0) continue;
        const int g = L.group_only >= 0 ? L.group_only : (blk & 7);
        const int jb = L.group_only >= 0 ? blk : (blk >> 3);
        // a block with one chunk of pixels per thread keeps their vertex and
        // normal in registers for the level (they do not depend on the pose)
        float4 cn[kIcpPx], cv[kIcpPx];
        const bool cached = nblk >= nb && blk < nb && icp_block_pixels(L, g, jb) <= kIcpPx;  // uniform
        if (cached) icp_load_pixels(L, g, jb, cn, cv);
        for (int it = 0; it < a.iters[level]; ++it) {
            float T[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) T[i] = Ts[i];
            if (solo) {
                // block 0 alone: the row stays in LDS
                icp_block_partials(L, T, g, jb, solo_rows + (8 * jb + g) * kPartStride, wsum);
                icp_group_sums(solo_rows, 8 * L.rows_step, L.group_only, wgsum, gsum);
            } else {
                // rows alternate between two buffers: a block writes exchange
                // n + 2 only after reading every row of n + 1, which every
                // block wrote after reading all of n
                unsigned long long *rows = a.rows[xchg & 1u];
                const unsigned tag = (epoch << 8) + xchg + 1u;
                unsigned long long *mine = rows + (8 * jb + g) * kPartStride;
                if (cached) {
                    icp_block_partials<true>(L, T, g, jb, nullptr, wsum, cn, cv, mine, tag);
                } else {
                    for (int v = blk; v < nb; v += nblk) {
                        const int gv = L.group_only >= 0 ? L.group_only : (v & 7);
                        const int jv = L.group_only >= 0 ? v : (v >> 3);
                        icp_block_partials(L, T, gv, jv, nullptr, wsum, nullptr, nullptr,
                                           rows + (8 * jv + gv) * kPartStride, tag);
                    }
                }
                if (!icp_group_sums_tagged(rows, 8 * L.rows_step, L.group_only, tag, a.spin_limit, &st->icp_abort,
                                           wgsum, gsum)) {
                    aborted = true;
                    break;
                }
                ++xchg;
            }
            if (blk == 0) icp_expose(L, gsum, st->partials);
            if (tid < 64) {
                double x[6] = {0, 0, 0, 0, 0, 0};
                float ae = 0.0f;
                const double nrm = icp_solve(L, gsum, solve_ws, x, ae);
                const double my_ang = lane == 1 ? x[4] : lane == 2 ? x[5] : x[3];
                double c1, s1;
                icp_sincos(my_ang, L.fast_trig != 0, c1, s1);
                const double cs[3] = {c1, __shfl(c1, 1, 64), __shfl(c1, 2, 64)};
                const double sn[3] = {s1, __shfl(s1, 1, 64), __shfl(s1, 2, 64)};
                if (lane == 0) {
                    float P[16];
#pragma unroll
                    for (int i = 0; i < 16; ++i) P[i] = Ts[i];
                    if (L.pose_fp64) {
                        double Pd[16];
                        for (int i = 0; i < 16; ++i) Pd[i] = Tsd[i];
                        pose_update_cs_d(x, cs, sn, P, Pd);
                        for (int i = 0; i < 16; ++i) Tsd[i] = Pd[i];
                    } else {
                        pose_update_cs(x, cs, sn, P);
                    }
#pragma unroll
                    for (int i = 0; i < 16; ++i) Ts[i] = P[i];
                    sh_done = (nrm < 1e-5 && nrm != 0.0) ? 1 : 0;
                    if (blk == 0) {
#pragma unroll
                        for (int k = 0; k < 6; ++k) st->x[k] = x[k];
                        st->ae = ae;
                    }
                }
            }
            __syncthreads();
            ++iters_run;
            if (sh_done) break;  // uniform over the grid
        }
    }
    if (blk == 0) {
        if (!aborted && tid < 16) {
            st->pose[tid] = Ts[tid];
            st->pose_d[tid] = Tsd[tid];
        }
        if (tid == 0) {
            st->iters_run = iters_run;
            st->level_done = sh_done;
            if (aborted) {
                // the pose is left as it was; nothing is integrated
                st->icp_error = 1;
                st->ae = 0.0f;
            }
        }
        __syncthreads();
        if (tid < 64) icp_finish_wave(st, a.frame, a.lv[0].library, a.lv[0].lib_max);
    }
    grid_exit(&st->icp_bar, &st->icp_exit, &st->icp_abort, (unsigned)nblk, &st->icp_epoch);
}

__global__ __launch_bounds__(kIcpThreads) void k_icp_track(IcpTrackArgs a, DevState *__restrict__ st)
{
    icp_track_body(a, st, blockIdx.x, gridDim.x);
}

// flags: kStateSetPose, kStateArmWindow (the 4-frame integrate window of
// kinect.cpp:146, 677-682), kStateResetPose (resetPose, gFusion.cpp:275-286:
// m_pose_eig = identity rotation with the pose's translation, m_cumTwist = 0,
// and the pose library cleared as Reset does at :238)
__global__ void k_state_set(DevState *st, Mat16 pose, int flags)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (flags & kStateSetPose)
        for (int i = 0; i < 16; ++i) st->pose[i] = pose.m[i];
    if (flags & kStateArmWindow) {
        st->counter = 0;
        st->reset = 1;
        st->integrate_now = 0;
        st->icp_error = 0;
    }
    if (flags & kStateResetPose) {
        for (int i = 0; i < 16; ++i) st->pose_eig[i] = (i % 5 == 0) ? 1.0f : 0.0f;
        st->pose_eig[12] = pose.m[12];
        st->pose_eig[13] = pose.m[13];
        st->pose_eig[14] = pose.m[14];
        for (int k = 0; k < 6; ++k) st->cum_twist[k] = 0.0;
        st->lib_n = 0;
        st->lib_overflow = 0;
        st->sdf_tracked = 0;
    }
    st->ticket = 0u;
    st->level_done = 0;
}

// ---------------------------------------------------------------------------
// Integrate
// ---------------------------------------------------------------------------
constexpr int kIntU = 8;      // z voxels per pipelined chunk (= brick depth)

struct ColumnSetup {
    float pos[3], cx[3], dl[3], cd[3];
};

__device__ __forceinline__ void column_setup(const IntegrateArgs &a, const float *pose, int x, int y, ColumnSetup &c)
{
    float inv[16], K[16];
    inverse4(pose, inv);  // inverse(m_pose) (gFusion.cpp:1242)
    camera_matrix(a.cam, K);
    const float vz = a.dim[2] / (float)a.N[2];
    // getVolumePosition(pix) with pix.z = 0 (integrate.cs:16-19, 57)
    const float c0 = ((float)x + 0.5f) * a.dim[0] / (float)a.N[0];
    const float c1 = ((float)y + 0.5f) * a.dim[1] / (float)a.N[1];
    const float c2 = (0.0f + 0.5f) * a.dim[2] / (float)a.N[2];
    opmul3(inv, c0, c1, c2, c.pos);
    opmul3(K, c.pos[0], c.pos[1], c.pos[2], c.cx);
    rotate3(inv, 0.0f, 0.0f, vz, c.dl);
    rotate3(K, c.dl[0], c.dl[1], c.dl[2], c.cd);
}

__device__ __forceinline__ void column_step(ColumnSetup &c)
{
    c.pos[0] += c.dl[0]; c.pos[1] += c.dl[1]; c.pos[2] += c.dl[2];
    c.cx[0] += c.cd[0]; c.cx[1] += c.cd[1]; c.cx[2] += c.cd[2];
}

// Projected pixel coordinate num/den + 0.5 (integrate.cs:67).  FAST uses the
// hardware reciprocal and falls back to the correctly rounded quotient when
// the result lies within a few ulps of an integer, so the pixel index and the
// image-bounds decision are always those of the exact expression.
template <bool FAST>
__device__ __forceinline__ float proj_coord(float num, float den, float rden)
{
    if (!FAST) return num / den + 0.5f;
    float p = num * rden + 0.5f;
    const float ap = fabsf(p);
    if (ap < 8192.0f) {
        const float fr = p - floorf(p);
        const float thr = fmaxf(ap * 6e-7f, 1e-30f);
        if (fr < thr || fr > 1.0f - thr) p = num / den + 0.5f;
    }
    return p;
}

// One thread per (x, y) z-column, 64 consecutive x per wavefront (coalesced
// plane accesses).  blockIdx.z splits the column into segments; a segment
// replays the reference's incremental stepping up to its start (ALU only) so
// every voxel sees exactly the positions of integrate.cs:62.  Each chunk of
// kIntU voxels first resolves projection + depth (loads batched), then issues
// the voxel loads it needs, then the updates and stores.  Every chunk that
// stores a negative value marks its 8^3 brick in the occupancy flags that the
// raycast uses to skip empty space (flags are only ever set here).
//
// FAST = false evaluates every division and square root correctly rounded, as
// the reference; FAST = true uses hardware reciprocal / square root (about 1
// ulp) for the signed distance and the weighted average.
template <class S, bool FAST>
__device__ __forceinline__ void integrate_body(const IntegrateArgs &a, const DevState *__restrict__ st, int bx, int by,
                                               int bz)
{
    if (a.gated && !st->integrate_now) return;
    const int x = bx * 64 + threadIdx.x;
    const int y = by * a.rows + threadIdx.y;
    if (x >= a.N[0] || y >= a.N[1]) return;

    ColumnSetup c;
    column_setup(a, st->pose, x, y, c);
    const int z0 = bz * a.zlen;
    const int z1 = min(z0 + a.zlen, a.N[2]);
    if (FAST) {
        // segment start evaluated directly (a few ulps from the stepped value)
        const float zf = (float)z0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            c.pos[k] = fmaf(zf, c.dl[k], c.pos[k]);
            c.cx[k] = fmaf(zf, c.cd[k], c.cx[k]);
        }
    } else {
        for (int z = 0; z < z0; ++z) column_step(c);  // the reference's stepping, replayed
    }
    typename S::V *__restrict__ tv = (typename S::V *)a.tsdf;
    typename S::V *__restrict__ wv = (typename S::V *)a.weight;
    const size_t plane = (size_t)a.N[0] * a.N[1];
    const size_t col = (size_t)y * a.N[0] + x;
    const float dmax_x = (float)(a.dw - 1), dmax_y = (float)(a.dh - 1);
    const float inv_mu = 1.0f / a.mu;
    uint8_t *__restrict__ bflags = a.bflags;
    const size_t brow = (size_t)(y >> 3) * a.NB[0] + (x >> 3);
    const size_t bplane = (size_t)a.NB[0] * a.NB[1];

    for (int zb = z0; zb < z1; zb += kIntU) {
        int pix[kIntU];
        float cz[kIntU], fac[kIntU];
#pragma unroll
        for (int k = 0; k < kIntU; ++k) {
            bool ok = (zb + k) < z1;
            ok = ok && !(c.pos[2] < 0.0001f);
            const float r = FAST ? __builtin_amdgcn_rcpf(c.cx[2]) : 0.0f;
            const float px = proj_coord<FAST>(c.cx[0], c.cx[2], r);
            const float py = proj_coord<FAST>(c.cx[1], c.cx[2], r);
            ok = ok && !(px < 0.0f || px > dmax_x || py < 0.0f || py > dmax_y);
            const int ix = (int)(ok ? px : 0.0f), iy = (int)(ok ? py : 0.0f);
            pix[k] = ok ? iy * a.dw + ix : -1;
            cz[k] = c.cx[2];
            if (FAST) {
                // pos.z == cameraX.z exactly (row 2 of K is (0, 0, 1, 0))
                const float qx = c.pos[0] * r, qy = c.pos[1] * r;
                fac[k] = __builtin_amdgcn_sqrtf((1.0f + qx * qx) + qy * qy);
            } else {
                const float qx = c.pos[0] / c.pos[2], qy = c.pos[1] / c.pos[2];
                fac[k] = sqrtf((1.0f + qx * qx) + qy * qy);
            }
            column_step(c);
        }
        float dep[kIntU];
#pragma unroll
        for (int k = 0; k < kIntU; ++k) dep[k] = a.depth_m[pix[k] < 0 ? 0 : pix[k]];
        int kind[kIntU];
        float sdf[kIntU];
#pragma unroll
        for (int k = 0; k < kIntU; ++k) {
            const float d = dep[k];  // depth / 1000 (integrate.cs:74), precomputed per pixel
            const float diff = (d - cz[k]) * fac[k];
            const bool live = pix[k] >= 0 && !(d == 0.0f);
            kind[k] = !live ? 0 : (diff > -a.mu ? 2 : 1);
            sdf[k] = gmin(1.0f, FAST ? diff * inv_mu : diff / a.mu);
        }
        // a brick with a clear zero flag holds only (0, 0): its voxels need not
        // be read, and writing (0, 0) to them changes nothing
        const size_t bidx = (size_t)(zb >> 3) * bplane + brow;
        const bool zf = a.zflags[bidx] != 0;
        typename S::V ot[kIntU], ow[kIntU];
#pragma unroll
        for (int k = 0; k < kIntU; ++k) {
            ot[k] = S::from_value(0.0f);
            ow[k] = S::from_weight(0.0f);
            if (kind[k] == 2 && zf) {
                const size_t vi = col + (size_t)(zb + k) * plane;
                ot[k] = tv[vi];
                ow[k] = wv[vi];
            }
        }
        bool neg = false, nz = false;
#pragma unroll
        for (int k = 0; k < kIntU; ++k) {
            if (kind[k] == 0) continue;
            const size_t vi = col + (size_t)(zb + k) * plane;
            float t = 0.0f, w = 0.0f;
            if (kind[k] == 2) {
                const float t0 = S::value(ot[k]), w0 = S::weight(ow[k]);
                const float wd = FAST ? (w0 * t0 + sdf[k]) * __builtin_amdgcn_rcpf(w0 + 1.0f)
                                      : (w0 * t0 + sdf[k]) / (w0 + 1.0f);
                if (wd < 0.2f) {
                    t = gmin(gmax(wd, -0.1f), 0.2f);
                    w = gmin(w0 + 1.0f, a.max_weight);
                }
            }
            const typename S::V nt = S::from_value(t), nw = S::from_weight(w);
            const bool zero = S::zero_bits(nt) && S::zero_bits(nw);
            if (zf || !zero) {
                tv[vi] = nt;
                wv[vi] = nw;
            }
            neg = neg || S::interesting(nt);
            nz = nz || !zero;
        }
        if (neg) bflags[bidx] = 1;
        if (nz && !zf) a.zflags[bidx] = 1;
    }
}

template <class S, bool FAST>
__global__ __launch_bounds__(256) void k_integrate(IntegrateArgs a, const DevState *__restrict__ st)
{
    integrate_body<S, FAST>(a, st, blockIdx.x, blockIdx.y, blockIdx.z);
}

// The voxels z of a column are at camera-frame positions pos + z * dl; the
// ones that can pass the image test of integrate.cs:64-70 satisfy five linear
// inequalities in z (depth in front of the near limit, pixel x and y inside
// the image).  Solving them in double with a 1-pixel and half-near-limit
// margin gives a range [zs, ze) outside of which every voxel of the column
// fails that test by more than the fast path's rounding, so only [zs, ze) is
// walked; inside it each voxel is still tested exactly.  zs stays on the
// segment's chunk grid (multiples of 8 from z0).
// column_frustum_range returns the whole column's range [floor(lo) - 1,
// ceil(hi) + 2) as integers clamped to +-2^30 (beyond every z, so the
// segment clip below compares as the unclamped values would); frustum_clip
// clips it to the segment [z0, z1) and aligns zs to the segment's chunk grid.
__device__ __forceinline__ int2 column_frustum_range(const ColumnSetup &c, float fx, float fy, float cxh, float cyh,
                                                     float dmax_x, float dmax_y)
{
    const double P0 = c.pos[0], P1 = c.pos[1], P2 = c.pos[2];
    const double D0 = c.dl[0], D1 = c.dl[1], D2 = c.dl[2];
    const double m = 1.0;
    double lo = -1e30, hi = 1e30;
    auto clip = [&](double A, double B) {  // keep z with A + B z >= 0
        if (B > 0.0)
            lo = fmax(lo, -A / B);
        else if (B < 0.0)
            hi = fmin(hi, -A / B);
        else if (A < 0.0)
            hi = -1e30;
    };
    clip(P2 - 0.00005, D2);
    clip((double)fx * P0 + ((double)cxh + m) * P2, (double)fx * D0 + ((double)cxh + m) * D2);
    clip(-((double)fx * P0 + ((double)cxh - (double)dmax_x - m) * P2),
         -((double)fx * D0 + ((double)cxh - (double)dmax_x - m) * D2));
    clip((double)fy * P1 + ((double)cyh + m) * P2, (double)fy * D1 + ((double)cyh + m) * D2);
    clip(-((double)fy * P1 + ((double)cyh - (double)dmax_y - m) * P2),
         -((double)fy * D1 + ((double)cyh - (double)dmax_y - m) * D2));
    // fmax / fmin ignore a NaN operand: a NaN bound clips nothing
    const double big = 1073741824.0;
    const double L = floor(lo) - 1.0, H = ceil(hi) + 2.0;
    return make_int2(L == L ? (int)fmin(fmax(L, -big), big) : -(1 << 30),
                     H == H ? (int)fmin(fmax(H, -big), big) : (1 << 30));
}

__device__ __forceinline__ void frustum_clip(int2 r, int z0, int z1, int &zs, int &ze)
{
    const int lo = max(r.x, z0), hi = min(r.y, z1);
    if (!(lo < hi)) {
        zs = ze = z0;
        return;
    }
    zs = z0 + ((lo - z0) / kIntU) * kIntU;
    ze = hi;
}

__device__ __forceinline__ void column_frustum(const ColumnSetup &c, float fx, float fy, float cxh, float cyh,
                                               float dmax_x, float dmax_y, int z0, int z1, int &zs, int &ze)
{
    frustum_clip(column_frustum_range(c, fx, fy, cxh, cyh, dmax_x, dmax_y), z0, z1, zs, ze);
}

// The per-column set-up of the fast integrate for one frame: one thread per
// (x, y) column computes what every z segment's thread of that column would
// (column_setup at z = 0, column_frustum_range), so the integrate threads
// load 24 bytes instead of inverting the pose and solving the frustum in
// double each; same values bit for bit.
__global__ __launch_bounds__(256) void k_integrate_columns(IntegrateArgs a, const DevState *__restrict__ st)
{
    if (a.gated && !st->integrate_now) return;
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= a.N[0] || y >= a.N[1]) return;
    ColumnSetup c;
    column_setup(a, st->pose, x, y, c);
    const float fx = a.cam[0], fy = a.cam[1], cxh = a.cam[2] + 0.5f, cyh = a.cam[3] + 0.5f;
    const size_t i = (size_t)y * a.N[0] + x;
    a.col_pos[i] = make_float4(c.pos[0], c.pos[1], c.pos[2], 0.0f);
    a.col_range[i] = column_frustum_range(c, fx, fy, cxh, cyh, (float)(a.dw - 1), (float)(a.dh - 1));
    if (x == 0 && y == 0) {
        a.col_dl[0] = c.dl[0];
        a.col_dl[1] = c.dl[1];
        a.col_dl[2] = c.dl[2];
    }
}

// In a brick holding only (0, 0) a voxel changes only if it is live with
// sdf < 0.2, i.e. depth - pos.z in (-mu, 0.2 mu) (diff = (depth - pos.z) * fac
// with fac >= 1; 1 % margins cover rounding).  The 8 voxels of a chunk lie on a
// segment from p to p + 7 dl; its projection stays inside the pixel box of the
// two end points (+-1 pixel), so if no nonzero depth of the 8x8 tiles covering
// that box lies within reach of the chunk's pos.z range, every voxel of the
// chunk keeps (0, 0) and the chunk needs no work.  Boxes wider than 2x2 tiles
// and chunks near the camera plane are always processed.
template <class Tiles>
__device__ __forceinline__ bool chunk_unreachable(const IntegrateArgs &a, const Tiles &tiles, const ColumnSetup &c,
                                                  float p0, float p1, float p2, float fx, float fy, float cxh,
                                                  float cyh, float mu_lo, float mu_hi)
{
    // straight-line: the four tile loads are always issued (at indices
    // clamped into the tile grid), so the tests of several chunks can be in
    // flight together; the result is that of the case analysis above
    const float q0 = fmaf(7.0f, c.dl[0], p0), q1 = fmaf(7.0f, c.dl[1], p1), q2 = fmaf(7.0f, c.dl[2], p2);
    const float zlo = fminf(p2, q2), zhi = fmaxf(p2, q2);
    const bool okz = zlo > 0.001f;
    const float rp = __builtin_amdgcn_rcpf(p2), rq = __builtin_amdgcn_rcpf(q2);
    const float xa = fmaf(fx, p0 * rp, cxh), xb = fmaf(fx, q0 * rq, cxh);
    const float ya = fmaf(fy, p1 * rp, cyh), yb = fmaf(fy, q1 * rq, cyh);
    const float W = (float)a.dw, H = (float)a.dh;
    const float x0 = fminf(xa, xb) - 1.0f, x1 = fmaxf(xa, xb) + 1.0f;
    const float y0 = fminf(ya, yb) - 1.0f, y1 = fmaxf(ya, yb) + 1.0f;
    const bool fin = x0 == x0 && x1 == x1 && y0 == y0 && y1 == y1;
    const bool off = x1 < 0.0f || y1 < 0.0f || x0 >= W || y0 >= H;  // whole box off the image
    // clamped into the image (med3 = min(max(., 0), limit) for the finite
    // values used; a NaN box is never used)
    const float cx0 = fin ? __builtin_amdgcn_fmed3f(x0, 0.0f, W - 1.0f) : 0.0f;
    const float cx1 = fin ? __builtin_amdgcn_fmed3f(x1, 0.0f, W - 1.0f) : 0.0f;
    const float cy0 = fin ? __builtin_amdgcn_fmed3f(y0, 0.0f, H - 1.0f) : 0.0f;
    const float cy1 = fin ? __builtin_amdgcn_fmed3f(y1, 0.0f, H - 1.0f) : 0.0f;
    const int tx0 = (int)cx0 >> 3, tx1 = (int)cx1 >> 3, ty0 = (int)cy0 >> 3, ty1 = (int)cy1 >> 3;
    const bool small = tx1 - tx0 <= 1 && ty1 - ty0 <= 1;
    const int txb = tx1 > tx0 ? tx0 + 1 : tx0, tyb = ty1 > ty0 ? ty0 + 1 : ty0;
    const float2 t00 = tiles(ty0 * a.tiles_x + tx0), t01 = tiles(ty0 * a.tiles_x + txb);
    const float2 t10 = tiles(tyb * a.tiles_x + tx0), t11 = tiles(tyb * a.tiles_x + txb);
    const float dmin = fminf(fminf(t00.x, t01.x), fminf(t10.x, t11.x));
    const float dmax = fmaxf(fmaxf(t00.y, t01.y), fmaxf(t10.y, t11.y));
    const float m = 1e-4f;
    const bool reach = zhi + m > dmin - mu_lo && zlo - m < dmax + mu_hi;
    return okz && fin && (off || (small && !reach));
}

// The fast integrate's reads of the small per-frame arrays: through buffer
// resources (32-bit offsets) when the volume's planes fit in 32-bit byte
// offsets, else plain global loads.
template <bool BUF>
struct IntegrateInputs;
template <>
struct IntegrateInputs<true> {
    __amdgpu_buffer_rsrc_t depth, tiles, zflags;
    __device__ __forceinline__ IntegrateInputs(const IntegrateArgs &a)
        : depth(__builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(a.depth_m), (short)0, a.dw * a.dh * 4,
                                                  0x00020000)),
          tiles(__builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(a.dtiles), (short)0,
                                                  a.tiles_x * a.tiles_y * 8, 0x00020000)),
          zflags(__builtin_amdgcn_make_buffer_rsrc(a.zflags, (short)0, a.NB[0] * a.NB[1] * a.NB[2], 0x00020000))
    {
    }
    // pixel index < 0 reads 0 (beyond the buffer); such values are not used
    __device__ __forceinline__ float depth_at(int pix) const
    {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(depth, pix * 4, 0, 0));
    }
    __device__ __forceinline__ float2 operator()(int t) const
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(tiles, t * 8, 0, 0);
        return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
    }
    __device__ __forceinline__ bool zflag(uint32_t b) const
    {
        return __builtin_amdgcn_raw_buffer_load_b8(zflags, (int)b, 0, 0) != 0;
    }
};
template <>
struct IntegrateInputs<false> {
    const float *d;
    const float2 *t;
    const uint8_t *z;
    __device__ __forceinline__ IntegrateInputs(const IntegrateArgs &a) : d(a.depth_m), t(a.dtiles), z(a.zflags) {}
    __device__ __forceinline__ float depth_at(int pix) const { return d[pix < 0 ? 0 : pix]; }
    __device__ __forceinline__ float2 operator()(int i) const { return t[i]; }
    __device__ __forceinline__ bool zflag(size_t b) const { return z[b] != 0; }
};

// Reduced-arithmetic integrate (KF_OPT_FAST_MATH).  Same update rule as
// k_integrate, with the per-voxel arithmetic cut to about a third: one
// hardware reciprocal of the camera depth per voxel, the projection folded to
// px = fx * (x / z) + cx + 0.5 (row 2 of K is (0, 0, 1, 0), so cameraX.z is
// pos.z), hardware square root, and reciprocals of mu and (w + 1).  Only the
// column's frustum range is walked (column_frustum); positions are stepped
// from its first chunk as in the reference.  Results agree with
// the correctly rounded evaluation to a few ulps; voxels on a pixel edge or a
// branch threshold can take the other side (counted as flips by the tests).
// OFF32: every byte offset into a TSDF plane fits in 32 bits; the planes,
// the depth image, the depth tiles and the zero flags are then read and
// written through buffer resources (one 32-bit offset per voxel, the z step
// of a chunk in the scalar offset).
// COUNT: the same walk, loads and decisions with no store and no flag
// update; counts[0] += voxels read, counts[1] += voxels written (the bytes
// this kernel moves for the volume are 2 x sizeof(V) per voxel of each).
template <class S, bool OFF32, bool COUNT = false>
__device__ __forceinline__ void integrate_fast_body(const IntegrateArgs &a, const DevState *__restrict__ st, int bx,
                                                    int by, int bz, unsigned long long *counts = nullptr)
{
    using Idx = typename std::conditional<OFF32, uint32_t, size_t>::type;
    const IntegrateInputs<OFF32> in(a);
    if (a.gated && !st->integrate_now) return;
    // columns beyond the volume take part with an empty range (the whole
    // wavefront takes part in the range reduction below)
    const bool col_in = bx * 64 + (int)threadIdx.x < a.N[0] && by * a.rows + (int)threadIdx.y < a.N[1];
    const int x = col_in ? bx * 64 + threadIdx.x : 0;
    const int y = col_in ? by * a.rows + threadIdx.y : 0;

    ColumnSetup c;
    const float fx = a.cam[0], fy = a.cam[1], cxh = a.cam[2] + 0.5f, cyh = a.cam[3] + 0.5f;
    const float dmax_x = (float)(a.dw - 1), dmax_y = (float)(a.dh - 1);
    const int z0 = bz * a.zlen;
    const int z1 = min(z0 + a.zlen, a.N[2]);
    int zs, ze;
    if (a.col_pos) {
        // the frame's per-column set-up (k_integrate_columns); the fast path
        // uses only the position at z = 0 and the z step
        const size_t ci = (size_t)y * a.N[0] + x;
        const float4 p = a.col_pos[ci];
        c.pos[0] = p.x;
        c.pos[1] = p.y;
        c.pos[2] = p.z;
        c.dl[0] = a.col_dl[0];
        c.dl[1] = a.col_dl[1];
        c.dl[2] = a.col_dl[2];
        frustum_clip(a.col_range[ci], z0, z1, zs, ze);
    } else {
        column_setup(a, st->pose, x, y, c);
        column_frustum(c, fx, fy, cxh, cyh, dmax_x, dmax_y, z0, z1, zs, ze);
    }
    if (!col_in) zs = ze = z0;
    // every lane tests the same chunks, from the wavefront's first chunk to
    // its last (zs is on the segment's chunk grid for every lane), so the
    // lanes of an iteration test one z: coalesced flag and tile loads and
    // neighbouring columns' similar skip decisions side by side; a lane's
    // chunks outside its own [zs, ze) are never needed
    int wzs = zs < ze ? zs : INT_MAX, wze = zs < ze ? ze : INT_MIN;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        wzs = min(wzs, __shfl_xor(wzs, o, 64));
        wze = max(wze, __shfl_xor(wze, o, 64));
    }
    if (wzs >= wze) return;  // the whole wavefront
    const float mu_lo = 0.2f * a.mu * 1.01f, mu_hi = a.mu * 1.01f;
    const float inv_mu = 1.0f / a.mu, neg_mu = -a.mu;
    typename S::V *__restrict__ tv = (typename S::V *)a.tsdf;
    typename S::V *__restrict__ wv = (typename S::V *)a.weight;
    // OFF32: the planes' byte offsets fit in 32 bits (buffer accesses)
    const uint32_t plane_bytes = OFF32 ? (uint32_t)((size_t)a.N[0] * a.N[1] * a.N[2] * sizeof(typename S::V)) : 0u;
    const BufPlane<S> TB(tv, plane_bytes), WB(wv, plane_bytes);
    const Idx plane = (Idx)a.N[0] * (Idx)a.N[1];
    const size_t bplane = (size_t)a.NB[0] * a.NB[1];
    const size_t brow_own = (size_t)(y >> 3) * a.NB[0] + (x >> 3);  // this lane's column of bricks
    unsigned long long n_read = 0, n_written = 0;  // COUNT only

    // windows of up to 16 chunks: first the skip tests of every chunk of the
    // window, one z at a time for the whole wavefront (so neighbouring
    // columns' tile and flag loads coincide) -- a brick with a clear zero
    // flag holds only (0, 0), so a chunk in one that no depth can reach keeps
    // (0, 0) and needs no work; then the needed chunks of all lanes as one
    // work list, 64 per pass (a surface crosses neighbouring columns at
    // different chunks and many columns cross none, so per-lane passes left
    // about half the lanes idle)
    constexpr int kWin = 16;  // chunks per window
    for (int wz = wzs; wz < wze; wz += kWin * kIntU) {
        const int wend = min(wze, wz + kWin * kIntU);
        uint32_t need = 0u, zfs = 0u;
        for (int i = 0, zb = wz; zb < wend; ++i, zb += kIntU) {
            const bool mine = zb >= zs && zb < ze;
            const bool zf = in.zflag((Idx)((zb >> 3) * (Idx)bplane + (Idx)brow_own));
            const bool unr = chunk_unreachable(a, in, c, fmaf((float)zb, c.dl[0], c.pos[0]),
                                               fmaf((float)zb, c.dl[1], c.pos[1]), fmaf((float)zb, c.dl[2], c.pos[2]),
                                               fx, fy, cxh, cyh, mu_lo, mu_hi);
            if (mine && !(unr && !zf)) need |= 1u << i;
            if (zf) zfs |= 1u << i;
        }
        // the window's needed chunks as one chunk-major work list: item
        // (lane, chunk i, brick-nonzero flag) at slot base_i + (rank of the
        // lane among the lanes needing chunk i); then 64 items per pass, each
        // lane taking the column of the item's lane, so every pass but the
        // last has all lanes at work and neighbouring lanes mostly hold
        // neighbouring columns at the same z
        __shared__ uint16_t work[4][kWin * 64];
        uint16_t *wl = work[threadIdx.y];
        const int lane = (int)threadIdx.x;
        const int nch = (wend - wz + kIntU - 1) / kIntU;
        uint32_t nitems = 0;
        for (int i = 0; i < nch; ++i) {
            const bool bit = (need >> i) & 1u;
            const unsigned long long bm = __ballot(bit);
            if (bit) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(bm >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
                wl[nitems + rank] = (uint16_t)(lane | (i << 6) | (((zfs >> i) & 1u) << 11));
            }
            nitems += (uint32_t)__popcll(bm);
        }
        __builtin_amdgcn_wave_barrier();
        const float cp0 = c.pos[0], cp1 = c.pos[1], cp2 = c.pos[2];
      for (uint32_t q0 = 0; q0 < nitems; q0 += 64) {
        const uint32_t q = q0 + (uint32_t)lane;
        const uint32_t it = q < nitems ? (uint32_t)wl[q] : (uint32_t)lane;
        const int sl = (int)(it & 63u), j = (int)((it >> 6) & 31u);
        // the item's column: its lane's position at z = 0 (all lanes take part)
        const float sp0 = __shfl(cp0, sl, 64), sp1 = __shfl(cp1, sl, 64), sp2 = __shfl(cp2, sl, 64);
        if (q >= nitems) continue;
        // (the wavefront's row: lanes beyond the volume hold y = 0)
        const int zb = wz + j * kIntU;
        const int sx = bx * 64 + sl, sy = by * a.rows + (int)threadIdx.y;
        const Idx col = (Idx)sy * (Idx)a.N[0] + (Idx)sx;
        const size_t brow = (size_t)(sy >> 3) * a.NB[0] + (sx >> 3);
        float p0 = fmaf((float)zb, c.dl[0], sp0);
        float p1 = fmaf((float)zb, c.dl[1], sp1);
        float p2 = fmaf((float)zb, c.dl[2], sp2);
        const size_t bidx = (size_t)(zb >> 3) * bplane + brow;
        const bool zf = (it >> 11) & 1u;
        int pix[kIntU];
        float pz[kIntU], fac[kIntU];
        const int kend = z1 - zb;  // voxels of the chunk inside the segment
#pragma unroll
        for (int k = 0; k < kIntU; ++k) {
            const float r = __builtin_amdgcn_rcpf(p2);
            const float qx = p0 * r, qy = p1 * r;
            const float px = fmaf(fx, qx, cxh), py = fmaf(fy, qy, cyh);
            // px in [0, dmax_x] exactly when clamping leaves it unchanged
            // (false for NaN); bitwise & keeps the predicate branch-free
            const float cpx = __builtin_amdgcn_fmed3f(px, 0.0f, dmax_x);
            const float cpy = __builtin_amdgcn_fmed3f(py, 0.0f, dmax_y);
            const bool ok = (k < kend) & !(p2 < 0.0001f) & (cpx == px) & (cpy == py);
            // 24-bit product: pixel indices < 2^24
            const int pv = (int)__umul24((uint32_t)(int)cpy, (uint32_t)a.dw) + (int)cpx;
            pix[k] = ok ? pv : -1;
            pz[k] = p2;
            fac[k] = __builtin_amdgcn_sqrtf(fmaf(qx, qx, fmaf(qy, qy, 1.0f)));
            p0 += c.dl[0];
            p1 += c.dl[1];
            p2 += c.dl[2];
        }
        float dep[kIntU];
#pragma unroll
        for (int k = 0; k < kIntU; ++k) dep[k] = in.depth_at(pix[k]);
        int kind[kIntU];
        float sdf[kIntU];
#pragma unroll
        for (int k = 0; k < kIntU; ++k) {
            const float diff = (dep[k] - pz[k]) * fac[k];
            const bool live = (pix[k] >= 0) & (dep[k] != 0.0f);
            kind[k] = !live ? 0 : (diff > neg_mu ? 2 : 1);
            sdf[k] = fminf(1.0f, diff * inv_mu);
        }
        // stored bits, zero-extended to 32 bits (no register packing of
        // 16-bit values, so the conditional loads stay independent and the
        // conversions wait for them only after all are issued)
        // voxel k of the chunk: element vc + k * plane; with OFF32 its byte
        // offset is voff + k * pbytes, the second term a scalar offset
        const Idx vc = col + (Idx)zb * plane;
        const uint32_t voff = OFF32 ? (uint32_t)vc * (uint32_t)sizeof(typename S::V) : 0u;
        const uint32_t pbytes = OFF32 ? (uint32_t)plane * (uint32_t)sizeof(typename S::V) : 0u;
        uint32_t ot[kIntU], ow[kIntU];
#pragma unroll
        for (int k = 0; k < kIntU; ++k) {
            ot[k] = 0u;
            ow[k] = 0u;
            if (kind[k] == 2 && zf) {
                const Idx vi = vc + (Idx)k * plane;
                if (OFF32) {
                    ot[k] = TB.bits(voff, (uint32_t)k * pbytes);
                    ow[k] = WB.bits(voff, (uint32_t)k * pbytes);
                } else {
                    ot[k] = S::bits(tv[vi]);
                    ow[k] = S::bits(wv[vi]);
                }
                if (COUNT) ++n_read;
            }
        }
        bool neg = false, nz = false;
#pragma unroll
        for (int k = 0; k < kIntU; ++k) {
            if (kind[k] == 0) continue;
            const Idx vi = vc + (Idx)k * plane;
            float t = 0.0f, w = 0.0f;
            if (kind[k] == 2) {
                const float t0 = S::value(S::from_bits(ot[k])), w0 = S::weight(S::from_bits(ow[k]));
                const float wd = fmaf(w0, t0, sdf[k]) * __builtin_amdgcn_rcpf(w0 + 1.0f);
                if (wd < 0.2f) {
                    t = gmax(wd, -0.1f);  // = min(max(wd, -0.1), 0.2) since wd < 0.2
                    w = fminf(w0 + 1.0f, a.max_weight);
                }
            }
            const typename S::V nt = S::from_value(t), nw = S::from_weight(w);
            const bool zero = S::zero_bits(nt) && S::zero_bits(nw);
            if (zf || !zero) {
                if (COUNT) {
                    ++n_written;
                } else if (OFF32) {
                    TB.store(voff, (uint32_t)k * pbytes, nt);
                    WB.store(voff, (uint32_t)k * pbytes, nw);
                } else {
                    tv[vi] = nt;
                    wv[vi] = nw;
                }
            }
            neg = neg || S::interesting(nt);
            nz = nz || !zero;
        }
        if (!COUNT) {
            if (neg) a.bflags[bidx] = 1;
            if (nz && !zf) a.zflags[bidx] = 1;
        }
      }
        __builtin_amdgcn_wave_barrier();  // the list is read before the next window writes it
    }
    if (COUNT) {
        atomicAdd(&counts[0], n_read);
        atomicAdd(&counts[1], n_written);
    }
}

template <class S, bool OFF32>
#ifndef KF_INT_WAVES_PER_SIMD
#define KF_INT_WAVES_PER_SIMD 6
#endif
__global__ __launch_bounds__(256, KF_INT_WAVES_PER_SIMD) void k_integrate_fast(IntegrateArgs a,
                                                                              const DevState *__restrict__ st)
{
    integrate_fast_body<S, OFF32>(a, st, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <class S, bool OFF32>
__global__ __launch_bounds__(256, KF_INT_WAVES_PER_SIMD) void k_count_integrate_fast(IntegrateArgs a,
                                                                                    const DevState *__restrict__ st,
                                                                                    unsigned long long *counts)
{
    integrate_fast_body<S, OFF32, true>(a, st, blockIdx.x, blockIdx.y, blockIdx.z, counts);
}

// Same walk as k_integrate without memory traffic on the volume: counts the
// voxels the reference semantics read (diff > -mu) and write (any live voxel).
__global__ __launch_bounds__(256) void k_count_integrate(IntegrateArgs a, const DevState *__restrict__ st,
                                                        unsigned long long *counts)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= a.N[0] || y >= a.N[1]) return;
    ColumnSetup c;
    column_setup(a, st->pose, x, y, c);
    unsigned long long nr = 0, nw = 0;
    const float dmax_x = (float)(a.dw - 1), dmax_y = (float)(a.dh - 1);
    for (int z = 0; z < a.N[2]; ++z) {
        bool ok = !(c.pos[2] < 0.0001f);
        const float px = c.cx[0] / c.cx[2] + 0.5f;
        const float py = c.cx[1] / c.cx[2] + 0.5f;
        ok = ok && !(px < 0.0f || px > dmax_x || py < 0.0f || py > dmax_y);
        if (ok) {
            const float d = a.depth_m[(int)py * a.dw + (int)px];
            if (!(d == 0.0f)) {
                const float qx = c.pos[0] / c.pos[2], qy = c.pos[1] / c.pos[2];
                const float diff = (d - c.cx[2]) * sqrtf((1.0f + qx * qx) + qy * qy);
                nw += 1;
                if (diff > -a.mu) nr += 1;
            }
        }
        column_step(c);
    }
    atomicAdd(&counts[0], nr);
    atomicAdd(&counts[1], nw);
}

// Exact brick flags: `flags` marks a brick any of whose voxels holds a value
// that can stop a ray (negative) or reach the 0.8 step switch; `zflags` marks
// a brick holding any voxel other than (0, 0).  A block covers 64 x 8 columns
// (8 bricks along x) and one brick layer in z; each wavefront reads one row of
// 64 consecutive voxels per z of both planes (coalesced), a ballot splits it
// into the 8 bricks, and LDS ORs the 8 rows.
// Incremental (a.incremental): between rebuilds every writer of the volume
// only sets flags (integrate sets both where it writes a negative / nonzero
// value; a reset clears volume and flags together), so a brick whose zero
// flag is clear holds only (0, 0) and keeps (0, 0) flags: only bricks with
// the zero flag set are read and rewritten.
template <class S>
__global__ __launch_bounds__(512) void k_brick_rebuild(BrickArgs a)
{
    __shared__ unsigned int any[16];
    if (threadIdx.x < 16) any[threadIdx.x] = 0u;
    const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane, y = blockIdx.y * 8 + row, bz = blockIdx.z;
    bool scan = x < a.N[0] && y < a.N[1];
    if (a.incremental && scan) scan = a.zflags[((size_t)bz * a.NB[1] + blockIdx.y) * a.NB[0] + (x >> 3)] != 0;
    // the zero flags are read before the block rewrites any
    __syncthreads();
    bool hit = false, nz = false;
    if (scan) {
        const typename S::V *v = (const typename S::V *)a.tsdf;
        const typename S::V *w = (const typename S::V *)a.weight;
        const size_t plane = (size_t)a.N[0] * a.N[1];
        const size_t base = (size_t)y * a.N[0] + x;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int z = bz * 8 + k;
            if (z < a.N[2]) {
                const typename S::V vv = v[(size_t)z * plane + base], ww = w[(size_t)z * plane + base];
                hit = hit || S::interesting(vv);
                nz = nz || !S::zero_bits(vv) || !S::zero_bits(ww);
            }
        }
    }
    const unsigned long long m = __ballot(hit), mz = __ballot(nz);
    if (lane < 8 && ((m >> (8 * lane)) & 0xFFull)) atomicOr(&any[lane], 1u);
    if (lane < 8 && ((mz >> (8 * lane)) & 0xFFull)) atomicOr(&any[8 + lane], 1u);
    __syncthreads();
    const int bx = blockIdx.x * 8 + (int)threadIdx.x;
    if (threadIdx.x < 8 && bx < a.NB[0]) {
        const size_t bi = ((size_t)bz * a.NB[1] + blockIdx.y) * a.NB[0] + bx;
        // (incremental: a brick with a clear zero flag was not read and
        // keeps its (0, 0) flags)
        if (!a.incremental || a.zflags[bi]) {
            a.flags[bi] = any[threadIdx.x] ? 1 : 0;
            a.zflags[bi] = any[8 + threadIdx.x] ? 1 : 0;
        }
    }
}

__global__ __launch_bounds__(256) void k_brick_dilate(BrickArgs a)
{
    brick_dilate_body(a, blockIdx.x * 4 + (threadIdx.x >> 6));
}

// ---------------------------------------------------------------------------
// Raycast
// ---------------------------------------------------------------------------
template <class S, bool TOUCH = false>
struct VolView {
    const typename S::V *__restrict__ v;
    int N0, N1, N2;
    size_t plane;
    uint32_t plane32;
    int off32;        // every byte offset into the value plane fits in 32 bits
    uint32_t *touch;  // TOUCH: one bit per voxel, set for every voxel read
    static constexpr uint32_t kSz = sizeof(typename S::V);
    __device__ __forceinline__ VolView(const RaycastArgs &a, uint32_t *touch_bits = nullptr)
        : v((const typename S::V *)a.tsdf), N0(a.N[0]), N1(a.N[1]), N2(a.N[2]), plane((size_t)a.N[0] * a.N[1]),
          plane32((uint32_t)((size_t)a.N[0] * a.N[1])), off32(a.off32), touch(touch_bits)
    {
    }
    __device__ __forceinline__ void mark(size_t i) const
    {
        if (TOUCH) atomicOr(&touch[i >> 5], 1u << (i & 31));
    }
    __device__ __forceinline__ float ld(size_t i) const
    {
        mark(i);
        return S::raw(v[i]);
    }
    // off32: the value plane is read at 32-bit byte offsets from uniform
    // base addresses (the base stays in scalar registers and each access
    // adds one 32-bit offset: no 64-bit address arithmetic per load)
    __device__ __forceinline__ float ld32(uint32_t i) const
    {
        if (TOUCH) return ld(i);
        return S::raw(*(const typename S::V *)((const char *)v + i * kSz));
    }
    // the two consecutive elements at byte offset voff from the uniform
    // address base (one load)
    struct alignas(kSz) Pair {
        typename S::V e0, e1;
    };
    __device__ __forceinline__ static void pair(const char *base, uint32_t voff, float &e0, float &e1)
    {
        const Pair q = *(const Pair *)(base + voff);
        e0 = S::raw(q.e0);
        e1 = S::raw(q.e1);
    }
    // the pair's stored bits as one dword (binary16 volumes: element x in
    // the low half, x + 1 in the high half)
    __device__ __forceinline__ static uint32_t pair_bits(const char *base, uint32_t voff)
    {
        static_assert(sizeof(Pair) == 4, "two 16-bit elements");
        const Pair q = *(const Pair *)(base + voff);
        return __builtin_bit_cast(uint32_t, q);
    }
    __device__ __forceinline__ float raw(int x, int y, int z) const
    {
        if ((unsigned)x >= (unsigned)N0 || (unsigned)y >= (unsigned)N1 || (unsigned)z >= (unsigned)N2) return 0.0f;
        return ld((size_t)z * plane + (size_t)y * N0 + x);
    }
};

// float -> int of an ivec3(floor(.)) cast; NaN / huge map out of bounds
__device__ __forceinline__ int floor_index(float fl)
{
    return (fl > -2147483648.0f && fl < 2147483648.0f) ? (int)fl : INT_MIN;
}

// interpVol: raycast.cs:69-82, split into the index setup (shared with the
// empty-space test) and the trilinear evaluation.
struct Sample {
    int lo[3], hi[3];
    float f[3], s[3];
    bool valid;
};

// p * N / dim - 0.5 on axis k.  When every dim is a power of two, 2^-m
// (the host allows only 2^-32 .. 2^32), the division is an exact scaling and
// p * N / dim = round(p N) 2^-m = round(p (N 2^-m)): rounding commutes with a
// power-of-two scaling while the result stays normal, and a result below the
// normal range (|p N 2^-m| < 2^-126) leaves s = -0.5 on either side, since
// the subtraction of 0.5 absorbs it.  So one multiply by the exact constant
// N 2^-m (vscale) gives the bits of the reference's multiply and divide.
// DP: 1 = every dim a power of two, 0 = not, -1 = read a.dim_pow2 (the
// raycast kernel is instantiated for both cases, so its hot loop carries no
// branch on it)
template <int DP = -1>
__device__ __forceinline__ float voxel_coord(const RaycastArgs &a, float p, int k)
{
    if (DP == 1 || (DP < 0 && a.dim_pow2)) return (p * a.vscale[k]) - 0.5f;
    return (p * (float)a.N[k] / a.dim[k]) - 0.5f;
}

template <int DP = -1>
__device__ __forceinline__ void sample_setup(const RaycastArgs &a, float p0, float p1, float p2, Sample &sm)
{
    sm.s[0] = voxel_coord<DP>(a, p0, 0);
    sm.s[1] = voxel_coord<DP>(a, p1, 1);
    sm.s[2] = voxel_coord<DP>(a, p2, 2);
    sm.valid = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float fl = floorf(sm.s[k]);
        const int b = floor_index(fl);
        sm.f[k] = sm.s[k] - fl;
        sm.valid = sm.valid && (b != INT_MIN);
        sm.lo[k] = (b == INT_MIN) ? b : max(b, 0);
        sm.hi[k] = (b == INT_MIN) ? b : min(b + 1, a.N[k] - 1);
    }
}

// interpVol's trilinear blend (raycast.cs:69-82), in its expression order
template <class S>
__device__ __forceinline__ float trilerp(float c000, float c100, float c010, float c110, float c001, float c101,
                                         float c011, float c111, const float *f)
{
    const float r = ((c000 * (1 - f[0]) + c100 * f[0]) * (1 - f[1]) + (c010 * (1 - f[0]) + c110 * f[0]) * f[1]) *
                        (1 - f[2]) +
                    ((c001 * (1 - f[0]) + c101 * f[0]) * (1 - f[1]) + (c011 * (1 - f[0]) + c111 * f[0]) * f[1]) *
                        f[2];
    return S::kScaled ? r * 0.00003051944088f : r;
}

// F16 volumes: the eight corner products of the blend read the stored halves
// directly.  v_fma_mix_f32 with a binary16 first source (the low or high
// half of a dword), an f32 weight and -0.0 as addend rounds the exact product
// once, sign of zero included: the bits of converting the half to f32 (exact)
// and multiplying, without the conversion instruction.
__device__ __forceinline__ float mul_half_lo(uint32_t h2, float w)
{
    float r;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h2), "v"(w), "v"(-0.0f));
    return r;
}
__device__ __forceinline__ float mul_half_hi(uint32_t h2, float w)
{
    float r;
    asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h2), "v"(w), "v"(-0.0f));
    return r;
}

// trilerp of four x-pairs of binary16 corners (low half x = b, high half
// x = b + 1): rows (y, z), (y + 1, z), (y, z + 1), (y + 1, z + 1); the same
// expression, in the same order, as trilerp<StoreF16>
__device__ __forceinline__ float trilerp_f16_pairs(const uint32_t (&r)[4], const float *f)
{
    const float a0 = mul_half_lo(r[0], 1 - f[0]) + mul_half_hi(r[0], f[0]);
    const float a1 = mul_half_lo(r[1], 1 - f[0]) + mul_half_hi(r[1], f[0]);
    const float a2 = mul_half_lo(r[2], 1 - f[0]) + mul_half_hi(r[2], f[0]);
    const float a3 = mul_half_lo(r[3], 1 - f[0]) + mul_half_hi(r[3], f[0]);
    return (a0 * (1 - f[1]) + a1 * f[1]) * (1 - f[2]) + (a2 * (1 - f[1]) + a3 * f[1]) * f[2];
}

template <class S, bool TOUCH>
__device__ __forceinline__ float sample_eval(const VolView<S, TOUCH> &V, const Sample &sm)
{
    const int *lo = sm.lo, *hi = sm.hi;
    const float *f = sm.f;
    float c000, c100, c010, c110, c001, c101, c011, c111;
    const bool inside = (unsigned)lo[0] < (unsigned)V.N0 && (unsigned)hi[0] < (unsigned)V.N0 &&
                        (unsigned)lo[1] < (unsigned)V.N1 && (unsigned)hi[1] < (unsigned)V.N1 &&
                        (unsigned)lo[2] < (unsigned)V.N2 && (unsigned)hi[2] < (unsigned)V.N2;
    if (V.off32 && inside) {
        // all 8 corners in the volume: 32-bit element offsets from one base
        // 24-bit products: lo, hi < N_k and plane32 = N0 N1 < 2^24 (off32)
        const uint32_t b = (__umul24((uint32_t)lo[2], V.plane32) + __umul24((uint32_t)lo[1], (uint32_t)V.N0)) +
                           (uint32_t)lo[0];
        const uint32_t dx = (uint32_t)(hi[0] - lo[0]);
        const uint32_t dy = __umul24((uint32_t)(hi[1] - lo[1]), (uint32_t)V.N0);
        const uint32_t dz = __umul24((uint32_t)(hi[2] - lo[2]), V.plane32);
        c000 = V.ld32(b);
        c100 = V.ld32(b + dx);
        c010 = V.ld32(b + dy);
        c110 = V.ld32(b + dy + dx);
        c001 = V.ld32(b + dz);
        c101 = V.ld32(b + dz + dx);
        c011 = V.ld32(b + dz + dy);
        c111 = V.ld32(b + dz + dy + dx);
    } else {  // corners may be out of bounds: they read 0 (image load semantics)
        c000 = V.raw(lo[0], lo[1], lo[2]);
        c100 = V.raw(hi[0], lo[1], lo[2]);
        c010 = V.raw(lo[0], hi[1], lo[2]);
        c110 = V.raw(hi[0], hi[1], lo[2]);
        c001 = V.raw(lo[0], lo[1], hi[2]);
        c101 = V.raw(hi[0], lo[1], hi[2]);
        c011 = V.raw(lo[0], hi[1], hi[2]);
        c111 = V.raw(hi[0], hi[1], hi[2]);
    }
    return trilerp<S>(c000, c100, c010, c110, c001, c101, c011, c111, f);
}

// Fine-step sample whose trilinear corners are b and b + 1 on every axis with
// no clamping (0 <= b_k <= N_k - 2): the index setup of sample_setup without
// the clamps.  Its eight corners are four pairs of x-neighbours, read as pairs
// from four row pointers; the values and the blend are those of
// sample_setup + sample_eval.
struct FastSample {
    uint32_t base;  // voxel index of corner (b0, b1, b2)
    int b[3];       // floor(s), INT_MIN when not a finite in-range value
    float f[3];
    bool inside;    // 0 <= b_k <= N_k - 2 on every axis
};

template <int DP = -1>
__device__ __forceinline__ void sample_fast(const RaycastArgs &a, float p0, float p1, float p2, uint32_t plane32,
                                            FastSample &q)
{
    const float p[3] = {p0, p1, p2};
    bool in = true;
    uint32_t bi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float sk = voxel_coord<DP>(a, p[k], k);
        const float fl = floorf(sk);
        q.b[k] = floor_index(fl);
        q.f[k] = sk - fl;
        // 0 <= b <= N - 2, tested on the integral float (false for NaN)
        in = in && fl >= 0.0f && fl <= (float)(a.N[k] - 2);
        bi[k] = (uint32_t)__float2int_rz(fl);  // = b when inside
    }
    q.inside = in;
    // 24-bit products: the base is used only when the sample is inside, where
    // b_k < N_k, and only with off32, where plane32 = N0 N1 < 2^24
    q.base = (__umul24(bi[2], plane32) + __umul24(bi[1], (uint32_t)a.N[0])) + bi[0];
}


template <class S, bool TOUCH, int DP = -1>
__device__ __forceinline__ float interp_vol(const VolView<S, TOUCH> &V, const RaycastArgs &a, float p0, float p1, float p2)
{
    Sample sm;
    sample_setup<DP>(a, p0, p1, p2, sm);
    return sample_eval(V, sm);
}

// getGradient: raycast.cs:229-276.  Stencil indices per axis are named
// kLL = b - 1, kLU = b (the trilinear "lo"), kUL = b + 1 ("hi"), kUU = b + 2,
// each clamped to the volume; VR(X, Y, Z) reads the voxel at those indices.
enum { kLL = 0, kLU = 1, kUL = 2, kUU = 3 };
#define KF_GRADIENT(VR, fx, fy, fz, g)                                                                              \
    g[0] = (((VR(kUL, kLU, kLU) - VR(kLL, kLU, kLU)) * (1 - fx) + (VR(kUU, kLU, kLU) - VR(kLU, kLU, kLU)) * fx) *    \
                (1 - fy) +                                                                                          \
            ((VR(kUL, kUL, kLU) - VR(kLL, kUL, kLU)) * (1 - fx) + (VR(kUU, kUL, kLU) - VR(kLU, kUL, kLU)) * fx) *    \
                fy) *                                                                                               \
               (1 - fz) +                                                                                           \
           (((VR(kUL, kLU, kUL) - VR(kLL, kLU, kUL)) * (1 - fx) + (VR(kUU, kLU, kUL) - VR(kLU, kLU, kUL)) * fx) *    \
                (1 - fy) +                                                                                          \
            ((VR(kUL, kUL, kUL) - VR(kLL, kUL, kUL)) * (1 - fx) + (VR(kUU, kUL, kUL) - VR(kLU, kUL, kUL)) * fx) *    \
                fy) *                                                                                               \
               fz;                                                                                                  \
    g[1] = (((VR(kLU, kUL, kLU) - VR(kLU, kLL, kLU)) * (1 - fx) + (VR(kUL, kUL, kLU) - VR(kUL, kLL, kLU)) * fx) *    \
                (1 - fy) +                                                                                          \
            ((VR(kLU, kUU, kLU) - VR(kLU, kLU, kLU)) * (1 - fx) + (VR(kUL, kUU, kLU) - VR(kUL, kLU, kLU)) * fx) *    \
                fy) *                                                                                               \
               (1 - fz) +                                                                                           \
           (((VR(kLU, kUL, kUL) - VR(kLU, kLL, kUL)) * (1 - fx) + (VR(kUL, kUL, kUL) - VR(kUL, kLL, kUL)) * fx) *    \
                (1 - fy) +                                                                                          \
            ((VR(kLU, kUU, kUL) - VR(kLU, kLU, kUL)) * (1 - fx) + (VR(kUL, kUU, kUL) - VR(kUL, kLU, kUL)) * fx) *    \
                fy) *                                                                                               \
               fz;                                                                                                  \
    g[2] = (((VR(kLU, kLU, kUL) - VR(kLU, kLU, kLL)) * (1 - fx) + (VR(kUL, kLU, kUL) - VR(kUL, kLU, kLL)) * fx) *    \
                (1 - fy) +                                                                                          \
            ((VR(kLU, kUL, kUL) - VR(kLU, kUL, kLL)) * (1 - fx) + (VR(kUL, kUL, kUL) - VR(kUL, kUL, kLL)) * fx) *    \
                fy) *                                                                                               \
               (1 - fz) +                                                                                           \
           (((VR(kLU, kLU, kUU) - VR(kLU, kLU, kLU)) * (1 - fx) + (VR(kUL, kLU, kUU) - VR(kUL, kLU, kLU)) * fx) *    \
                (1 - fy) +                                                                                          \
            ((VR(kLU, kUL, kUU) - VR(kLU, kUL, kLU)) * (1 - fx) + (VR(kUL, kUL, kUU) - VR(kUL, kUL, kLU)) * fx) *    \
                fy) *                                                                                               \
               fz;

template <class S, bool TOUCH>
__device__ void gradient(const VolView<S, TOUCH> &V, const RaycastArgs &a, const float hit[3], float g[3])
{
    int ix[3][4];
    int bb[3];
    float f[3];
    bool interior = !TOUCH && a.off32;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float s = (hit[k] * (float)a.N[k] / a.dim[k]) - 0.5f;
        const float fl = floorf(s);
        const int b = floor_index(fl);
        f[k] = s - fl;
        bb[k] = b;
        if (b == INT_MIN) {
            ix[k][kLL] = ix[k][kLU] = ix[k][kUL] = ix[k][kUU] = b;
        } else {
            ix[k][kLL] = max(b - 1, 0);
            ix[k][kLU] = max(b, 0);
            ix[k][kUL] = min(b + 1, a.N[k] - 1);
            ix[k][kUU] = min(b + 2, a.N[k] - 1);
        }
        // every stencil index in the volume without clamping: 1 <= b <= N - 3
        interior = interior && ((unsigned)b - 1u < (unsigned)max(a.N[k] - 3, 0));
    }
    const float fx = f[0], fy = f[1], fz = f[2];
    if (interior) {
        // the 32 voxels of the stencil from twelve row reads: x = b - 1 .. b + 2
        // at y, z in {b, b + 1}; x in {b, b + 1} at y in {b - 1, b + 2} and
        // at z in {b - 1, b + 2}.  Byte offsets from the corner (b - 1, b - 1,
        // b - 1) (interior: b >= 1); each row's (y, z) displacement is folded
        // into a uniform base address.
        constexpr uint32_t sz = VolView<S, TOUCH>::kSz;
        const uint32_t corner = (((uint32_t)bb[2] - 1u) * V.plane32 + ((uint32_t)bb[1] - 1u) * (uint32_t)V.N0) +
                                ((uint32_t)bb[0] - 1u);
        const uint32_t vo = corner * sz;
        const size_t rowb = (size_t)V.N0 * sz, plb = (size_t)V.plane32 * sz;
        const char *v0 = (const char *)V.v;
        auto row = [&](int yi, int zi) { return v0 + (size_t)yi * rowb + (size_t)zi * plb; };
        float A[2][2][4], B[2][2][2], C[2][2][2];
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int z = 0; z < 2; ++z) {
                const typename S::V *r = (const typename S::V *)(row(y + 1, z + 1) + vo);
#pragma unroll
                for (int x = 0; x < 4; ++x) A[y][z][x] = S::raw(r[x]);
            }
#pragma unroll
        for (int z = 0; z < 2; ++z)
#pragma unroll
            for (int y = 0; y < 2; ++y) V.pair(row(y ? 3 : 0, z + 1), vo + sz, B[z][y][0], B[z][y][1]);
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int z = 0; z < 2; ++z) V.pair(row(y + 1, z ? 3 : 0), vo + sz, C[y][z][0], C[y][z][1]);
        auto VF = [&](int X, int Y, int Z) -> float {
            if ((Y == kLU || Y == kUL) && (Z == kLU || Z == kUL)) return A[Y - kLU][Z - kLU][X];
            if (Y == kLL || Y == kUU) return B[Z - kLU][Y == kLL ? 0 : 1][X - kLU];
            return C[Y - kLU][Z == kLL ? 0 : 1][X - kLU];
        };
        KF_GRADIENT(VF, fx, fy, fz, g)
    } else {
        auto VR = [&](int X, int Y, int Z) -> float { return V.raw(ix[0][X], ix[1][Y], ix[2][Z]); };
        KF_GRADIENT(VR, fx, fy, fz, g)
    }
    const float gs = S::kScaled ? (0.5f * 0.00003051944088f) : 0.5f;
#pragma unroll
    for (int k = 0; k < 3; ++k) g[k] = (g[k] * (a.dim[k] / (float)a.N[k])) * gs;
}

constexpr int kMaxMarchSteps = 1 << 20;
#ifndef KF_RAYCAST_BATCH
#define KF_RAYCAST_BATCH 4
#endif
constexpr int kBatch = KF_RAYCAST_BATCH;  // fine-step samples evaluated per march iteration

// view = m_pose * invK (gFusion.cpp:1298)
__device__ __forceinline__ void raycast_view(const RaycastArgs &a, const DevState *__restrict__ st, float view[16])
{
    float iK[16], P[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) P[i] = st->pose[i];
    inverse_camera_matrix(a.cam, iK);
    mul4(P, iK, view);
}

// The ray of pixel (x, y): march, hit refinement, normal, map writes.
// Returns the ray's loop iterations counted as empty cells entered +
// fine-step batches of four samples.
template <class S, bool TOUCH, int DP>
__device__ __forceinline__ unsigned raycast_pixel(const RaycastArgs &a, const float (&view)[16], int x, int y)
{
    const VolView<S, TOUCH> V(a, a.touch);

    // TOUCH: the flag bytes read are marked after the voxel bits (dilated
    // brick flags, then coarse cells)
    const size_t nvox = (size_t)a.N[0] * a.N[1] * a.N[2], nbt = (size_t)a.NB[0] * a.NB[1] * a.NB[2];
    auto flag_mark = [&](int bi, int ci) {
        V.mark(nvox + (size_t)bi);
        V.mark(nvox + nbt + (size_t)ci);
    };
    (void)flag_mark;
    const float o[3] = {view[12], view[13], view[14]};
    float d[3];
    rotate3(view, (float)x, (float)y, 1.0f, d);
    float tmin[3], tmax[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float invR = 1.0f / d[k];
        const float tbot = (-1.0f * invR) * o[k];
        const float ttop = invR * (a.dim[k] - o[k]);
        tmin[k] = gmin(ttop, tbot);
        tmax[k] = gmax(ttop, tbot);
    }
    const float largest_tmin = gmax(gmax(tmin[0], tmin[1]), gmax(tmin[0], tmin[2]));
    const float smallest_tmax = gmin(gmin(tmax[0], tmax[1]), gmin(tmax[0], tmax[2]));
    const float tnear = gmax(largest_tmin, a.near_plane);
    const float tfar = gmin(smallest_tmax, a.far_plane);

    bool hit = false;
    float hp[3], th = 0.0f;
    unsigned n_eval = 0, n_skip = 0, n_brick = 0;  // statistics mode only
    if (tnear < tfar) {
        float t = tnear, stepsize = a.large_step;
        float f_t = interp_vol<S, TOUCH, DP>(V, a, o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t);
        float f_tt = 0.0f;
        if (f_t >= 0.0f) {
            // Empty-space skipping.  A sample whose lower trilinear corner lies
            // in a brick with a clear dilated flag reads only voxels in
            // [0, 0.79): its value is >= 0 and < 0.8, so it can neither end
            // the march nor change the step.  Such samples are not evaluated,
            // but t still advances by the same float additions, so every
            // evaluated sample sits at exactly the reference's t.
            const float ds[3] = {d[0] * (float)a.N[0] / a.dim[0], d[1] * (float)a.N[1] / a.dim[1],
                                 d[2] * (float)a.N[2] / a.dim[2]};
            // 1 / ds: the exit bound below is a skip limit with a 0.01-voxel
            // margin, far wider than the ulp a reciprocal adds
            const float rds[3] = {1.0f / ds[0], 1.0f / ds[1], 1.0f / ds[2]};
            float t_prev = t;
            bool have_prev = true;
            n_eval++;
            int guard = 0;
            // flags of the 32^3 cell and 8^3 brick of the next sample, loaded
            // together with the previous fine-step batch (have_flags)
            uint8_t cf_n = 1, ff_n = 1;
            bool have_flags = false;
            for (; t < tfar && guard < kMaxMarchSteps; ++guard) {
                const bool can_skip = a.skip && stepsize == a.step;
                Sample sm;
                sample_setup<DP>(a, o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t, sm);
                if (can_skip && sm.valid) {
                    // two-level occupancy: 32^3 coarse cells, then 8^3 bricks
                    uint8_t cf = cf_n, ff = ff_n;
                    if (!have_flags) {
                        const int b0 = sm.lo[0] >> 3, b1 = sm.lo[1] >> 3, b2 = sm.lo[2] >> 3;
                        cf = ff = 1;  // outside the brick grid: never skipped
                        if (b0 < a.NB[0] && b1 < a.NB[1] && b2 < a.NB[2]) {
                            const int ci = ((b2 >> 2) * a.NC[1] + (b1 >> 2)) * a.NC[0] + (b0 >> 2);
                            const int bi = (b2 * a.NB[1] + b1) * a.NB[0] + b0;
                            cf = a.cflags[ci];
                            ff = a.dflags[bi];
                            flag_mark(bi, ci);
                        }
                    }
                    const int shift = cf == 0 ? 5 : (ff == 0 ? 3 : 0);
                    if (shift) {
                        // t below which the lower corner stays in the empty cell
                        // (0.01-voxel margin covers the rounding of p and s)
                        const float span = (float)(1 << shift);
                        float ts = INFINITY;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const float base = (float)((sm.lo[k] >> shift) << shift);
                            if (ds[k] > 0.0f)
                                ts = fminf(ts, t + (base + span - 0.01f - sm.s[k]) * rds[k]);
                            else if (ds[k] < 0.0f && base > 0.0f)
                                ts = fminf(ts, t + (base + 0.01f - sm.s[k]) * rds[k]);
                        }
                        // this sample and every later one below ts: same float steps as the reference
                        const int ns = march_to(t, t_prev, stepsize, fminf(ts, tfar));
                        n_skip += ns;
                        guard += ns;
                        have_prev = false;
                        have_flags = false;
                        n_brick++;
                        continue;
                    }
                }
                have_flags = false;
                if (stepsize != a.step) {
                    // still on the large step: one sample at a time (raycast.cs:185-212)
                    f_tt = sample_eval(V, sm);
                    n_eval++;
                    if (f_tt < 0.0f) break;
                    if (f_tt < 0.8f) stepsize = a.step;
                    f_t = f_tt;
                    have_prev = true;
                    t_prev = t;
                    t += stepsize;
                    continue;
                }
                // fine step: evaluate kBatch consecutive samples at once (their
                // loads are independent) and fetch the flags of the sample after
                // them, then consume the batch in march order.
                float tj[kBatch], fv[kBatch];
                tj[0] = t;
#pragma unroll
                for (int j = 1; j < kBatch; ++j) tj[j] = tj[j - 1] + stepsize;
                // interior samples: corners b and b + 1 on every axis
                FastSample fq[kBatch];
                bool all_in = !TOUCH && a.off32;
                if (all_in) {
                    fq[0].inside = sm.valid;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        fq[0].b[k] = sm.lo[k];
                        fq[0].f[k] = sm.f[k];
                        fq[0].inside = fq[0].inside && sm.lo[k] >= 0 && sm.hi[k] == sm.lo[k] + 1;
                    }
                    // 24-bit products, as in sample_fast (used only when inside)
                    fq[0].base = (__umul24((uint32_t)sm.lo[2], V.plane32) + __umul24((uint32_t)sm.lo[1], (uint32_t)a.N[0])) +
                                 (uint32_t)sm.lo[0];
#pragma unroll
                    for (int j = 1; j < kBatch; ++j)
                        sample_fast<DP>(a, o[0] + d[0] * tj[j], o[1] + d[1] * tj[j], o[2] + d[2] * tj[j], V.plane32,
                                    fq[j]);
#pragma unroll
                    for (int j = 0; j < kBatch; ++j) all_in = all_in && fq[j].inside;
                }
                if (__all(all_in)) {
                    // every active lane: all loads first, then the blends
                    constexpr uint32_t sz = VolView<S, TOUCH>::kSz;
                    const size_t rowb = (size_t)V.N0 * sz, plb = (size_t)V.plane32 * sz;
                    // the four row bases are uniform; each sample adds one
                    // 32-bit byte offset
                    const char *rb0 = (const char *)V.v;
                    const char *rb1 = rb0 + rowb, *rb2 = rb0 + plb, *rb3 = rb0 + (plb + rowb);
                    if constexpr (std::is_same<S, StoreF16>::value) {
                        // binary16: the pairs stay packed, the blend reads the halves
                        uint32_t r[kBatch][4];
#pragma unroll
                        for (int j = 0; j < kBatch; ++j) {
                            const uint32_t vo = fq[j].base * sz;
                            r[j][0] = V.pair_bits(rb0, vo);
                            r[j][1] = V.pair_bits(rb1, vo);
                            r[j][2] = V.pair_bits(rb2, vo);
                            r[j][3] = V.pair_bits(rb3, vo);
                        }
#pragma unroll
                        for (int j = 0; j < kBatch; ++j) fv[j] = trilerp_f16_pairs(r[j], fq[j].f);
                    } else {
                        float c[kBatch][8];
#pragma unroll
                        for (int j = 0; j < kBatch; ++j) {
                            // rows (y, z), (y + 1, z), (y, z + 1), (y + 1, z + 1)
                            const uint32_t vo = fq[j].base * sz;
                            V.pair(rb0, vo, c[j][0], c[j][1]);
                            V.pair(rb1, vo, c[j][2], c[j][3]);
                            V.pair(rb2, vo, c[j][4], c[j][5]);
                            V.pair(rb3, vo, c[j][6], c[j][7]);
                        }
#pragma unroll
                        for (int j = 0; j < kBatch; ++j)
                            fv[j] = trilerp<S>(c[j][0], c[j][1], c[j][2], c[j][3], c[j][4], c[j][5], c[j][6],
                                               c[j][7], fq[j].f);
                    }
                } else {
                    fv[0] = sample_eval(V, sm);
#pragma unroll
                    for (int j = 1; j < kBatch; ++j) {
                        Sample sj;
                        sample_setup<DP>(a, o[0] + d[0] * tj[j], o[1] + d[1] * tj[j], o[2] + d[2] * tj[j], sj);
                        fv[j] = sample_eval(V, sj);
                    }
                }
                const float t_next = tj[kBatch - 1] + stepsize;
                if (a.skip && t_next < tfar) {
                    // the next sample lies in the box: its lower corner is a
                    // voxel of the volume (checked all the same); the corner
                    // index is floor(s) clamped at 0, as in sample_setup
                    FastSample fn;
                    sample_fast<DP>(a, o[0] + d[0] * t_next, o[1] + d[1] * t_next, o[2] + d[2] * t_next, V.plane32, fn);
                    const bool valid = fn.b[0] != INT_MIN && fn.b[1] != INT_MIN && fn.b[2] != INT_MIN;
                    const int b0 = max(fn.b[0], 0) >> 3, b1 = max(fn.b[1], 0) >> 3, b2 = max(fn.b[2], 0) >> 3;
                    if (valid && b0 < a.NB[0] && b1 < a.NB[1] && b2 < a.NB[2]) {
                        const int ci = ((b2 >> 2) * a.NC[1] + (b1 >> 2)) * a.NC[0] + (b0 >> 2);
                        const int bi = (b2 * a.NB[1] + b1) * a.NB[0] + b0;
                        cf_n = a.cflags[ci];
                        ff_n = a.dflags[bi];
                        flag_mark(bi, ci);
                        have_flags = true;
                    }
                }
                // consumed in march order: m is the first sample that ends
                // the march (t past tfar, or a negative value), kBatch if
                // none; the samples before it are taken (f >= 0 and < 0.8
                // keep the fine step), the last of them becomes f_t / t_prev
                int m = kBatch;
#pragma unroll
                for (int j = kBatch - 1; j >= 0; --j)
                    if (!(tj[j] < tfar) || fv[j] < 0.0f) m = j;
                if (m > 0) {
                    float tp = tj[0], fp = fv[0];
#pragma unroll
                    for (int j = 1; j < kBatch; ++j)
                        if (m > j) {
                            tp = tj[j];
                            fp = fv[j];
                        }
                    t_prev = tp;
                    f_t = fp;
                    have_prev = true;
                }
                if (m < kBatch) {
                    float tm = tj[0], fm = fv[0];
#pragma unroll
                    for (int j = 1; j < kBatch; ++j)
                        if (m == j) {
                            tm = tj[j];
                            fm = fv[j];
                        }
                    t = tm;
                    if (tm < tfar) {  // a negative sample: the hit; else the march ended before it
                        f_tt = fm;
                        n_eval += m + 1;
                    } else {
                        n_eval += m;
                    }
                    break;
                }
                n_eval += kBatch;
                t = t_next;
                guard += kBatch - 1;
            }
            if (f_tt < 0.0f) {
                if (!have_prev) f_t = interp_vol<S, TOUCH, DP>(V, a, o[0] + d[0] * t_prev, o[1] + d[1] * t_prev, o[2] + d[2] * t_prev);
                t = t + stepsize * f_tt / (f_t - f_tt);
                hp[0] = o[0] + d[0] * t;
                hp[1] = o[1] + d[1] * t;
                hp[2] = o[2] + d[2] * t;
                th = t;
                hit = true;
            }
        }
    }
    const int i = y * a.w + x;
    if (hit && th > 0.0f) {
        a.refV[i] = make_float4(hp[0], hp[1], hp[2], 1.0f);
        float g[3];
        gradient(V, a, hp, g);
        const float len = sqrtf(dot3(g[0], g[1], g[2], g[0], g[1], g[2]));
        if (len == 0.0f)
            a.refN[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        else
            a.refN[i] = make_float4(g[0] / len, g[1] / len, g[2] / len, 1.0f);
    } else {
        a.refV[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.refN[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (a.stats) {
        atomicAdd(&a.stats[0], (unsigned long long)n_eval);
        atomicAdd(&a.stats[1], (unsigned long long)n_skip);
        atomicAdd(&a.stats[2], (unsigned long long)n_brick);
        atomicAdd(&a.stats[3], (unsigned long long)(hit && th > 0.0f));
        atomicMax(&a.stats[4], (unsigned long long)n_eval);
        atomicMax(&a.stats[5], (unsigned long long)(n_eval + n_brick));
        if (a.ray_counts) a.ray_counts[i] = min(n_eval, 0xffffu) | (min(n_brick, 0xffffu) << 16);
    }
    return n_brick + (n_eval + 3u) / 4u;
}

// The same march with four lanes per ray (8x8-pixel tiles, 64 rays per
// block).  The four lanes hold identical march state and make identical
// decisions (occupancy flags, fast-forwards, the large step); on the fine
// step they evaluate the next 16 samples together, lane q taking samples q,
// q + 4, q + 8, q + 12, each at the t the reference's additions give, and a
// two-step shuffle finds the first sample (in march order) that ends the
// march.  A ray's chain of dependent loads is four times shorter and the
// grid has four times as many wavefronts to hide latency.
constexpr int kCoopB = 4;                 // samples per lane per round
constexpr int kCoopN = 4 * kCoopB;        // consecutive samples per round

// The march of the ray of pixel (x, y) by a group of four lanes (lane q of
// the group, whose first lane in the wavefront is gbase); the group's lane 0
// writes the maps.  Returns the ray's iterations as the one-lane march counts
// them (empty cells entered + fine-step batches of four samples), the same
// on all four lanes.
template <class S, int DP = -1>
__device__ __forceinline__ unsigned raycast_coop_ray(const RaycastArgs &a, const float (&view)[16], int x, int y,
                                                     int q, int gbase)
{
    const VolView<S> V(a);

    const float o[3] = {view[12], view[13], view[14]};
    float d[3];
    rotate3(view, (float)x, (float)y, 1.0f, d);
    float tmin[3], tmax[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float invR = 1.0f / d[k];
        const float tbot = (-1.0f * invR) * o[k];
        const float ttop = invR * (a.dim[k] - o[k]);
        tmin[k] = gmin(ttop, tbot);
        tmax[k] = gmax(ttop, tbot);
    }
    const float largest_tmin = gmax(gmax(tmin[0], tmin[1]), gmax(tmin[0], tmin[2]));
    const float smallest_tmax = gmin(gmin(tmax[0], tmax[1]), gmin(tmax[0], tmax[2]));
    const float tnear = gmax(largest_tmin, a.near_plane);
    const float tfar = gmin(smallest_tmax, a.far_plane);

    bool hit = false;
    float hp[3], th = 0.0f;
    unsigned n_eval = 0, n_skip = 0, n_brick = 0;  // statistics mode only
    if (tnear < tfar) {
        float t = tnear, stepsize = a.large_step;
        float f_t = interp_vol<S, false, DP>(V, a, o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t);
        float f_tt = 0.0f;
        if (f_t >= 0.0f) {
            const float ds[3] = {d[0] * (float)a.N[0] / a.dim[0], d[1] * (float)a.N[1] / a.dim[1],
                                 d[2] * (float)a.N[2] / a.dim[2]};
            // 1 / ds: the exit bound below is a skip limit with a 0.01-voxel
            // margin, far wider than the ulp a reciprocal adds
            const float rds[3] = {1.0f / ds[0], 1.0f / ds[1], 1.0f / ds[2]};
            float t_prev = t;
            bool have_prev = true;
            n_eval++;
            int guard = 0;
            for (; t < tfar && guard < kMaxMarchSteps; ++guard) {
                const bool can_skip = a.skip && stepsize == a.step;
                Sample sm;
                sample_setup<DP>(a, o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t, sm);
                if (can_skip && sm.valid) {
                    // empty-space skipping exactly as the one-lane march
                    const int b0 = sm.lo[0] >> 3, b1 = sm.lo[1] >> 3, b2 = sm.lo[2] >> 3;
                    uint8_t cf = 1, ff = 1;  // outside the brick grid: never skipped
                    if (b0 < a.NB[0] && b1 < a.NB[1] && b2 < a.NB[2]) {
                        cf = a.cflags[((b2 >> 2) * a.NC[1] + (b1 >> 2)) * a.NC[0] + (b0 >> 2)];
                        ff = a.dflags[(b2 * a.NB[1] + b1) * a.NB[0] + b0];
                    }
                    const int shift = cf == 0 ? 5 : (ff == 0 ? 3 : 0);
                    if (shift) {
                        const float span = (float)(1 << shift);
                        float ts = INFINITY;
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const float base = (float)((sm.lo[k] >> shift) << shift);
                            if (ds[k] > 0.0f)
                                ts = fminf(ts, t + (base + span - 0.01f - sm.s[k]) * rds[k]);
                            else if (ds[k] < 0.0f && base > 0.0f)
                                ts = fminf(ts, t + (base + 0.01f - sm.s[k]) * rds[k]);
                        }
                        const int ns = march_to(t, t_prev, stepsize, fminf(ts, tfar));
                        n_skip += ns;
                        guard += ns;
                        have_prev = false;
                        n_brick++;
                        continue;
                    }
                }
                if (stepsize != a.step) {
                    // the large step: one sample (raycast.cs:185-212), the same on every lane
                    f_tt = sample_eval(V, sm);
                    n_eval++;
                    if (f_tt < 0.0f) break;
                    if (f_tt < 0.8f) stepsize = a.step;
                    f_t = f_tt;
                    have_prev = true;
                    t_prev = t;
                    t += stepsize;
                    continue;
                }
                // fine step: this lane's samples q, q + 4, q + 8, q + 12 of the
                // next 16, at the t of the reference's sequential additions
                float tb[kCoopB], fv[kCoopB];
                {
                    float tc = t;
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        if (k < q) tc += stepsize;
                    tb[0] = tc;
#pragma unroll
                    for (int b = 1; b < kCoopB; ++b) {
                        float u = tb[b - 1];
#pragma unroll
                        for (int k = 0; k < 4; ++k) u += stepsize;
                        tb[b] = u;
                    }
                }
#pragma unroll
                for (int b = 0; b < kCoopB; ++b) {
                    Sample sb;
                    sample_setup<DP>(a, o[0] + d[0] * tb[b], o[1] + d[1] * tb[b], o[2] + d[2] * tb[b], sb);
                    fv[b] = sample_eval(V, sb);
                }
                int mine = kCoopN;  // first sample of this lane that ends the march
#pragma unroll
                for (int b = kCoopB - 1; b >= 0; --b)
                    if (!(tb[b] < tfar) || fv[b] < 0.0f) mine = 4 * b + q;
                int m = min(mine, __shfl_xor(mine, 1, 64));
                m = min(m, __shfl_xor(m, 2, 64));
                if (m == kCoopN) {
                    // all 16 taken; the last (lane 3, slot 3) becomes f_t
                    f_t = __shfl(fv[kCoopB - 1], gbase + 3, 64);
                    t_prev = __shfl(tb[kCoopB - 1], gbase + 3, 64);
                    have_prev = true;
                    t = t_prev + stepsize;
                    n_eval += kCoopN;
                    guard += kCoopN - 1;
                    continue;
                }
                const int mb = m >> 2;
                const float tsel = mb == 0 ? tb[0] : mb == 1 ? tb[1] : mb == 2 ? tb[2] : tb[3];
                const float fsel = mb == 0 ? fv[0] : mb == 1 ? fv[1] : mb == 2 ? fv[2] : fv[3];
                const float tm = __shfl(tsel, gbase + (m & 3), 64);
                const float fm = __shfl(fsel, gbase + (m & 3), 64);
                if (m > 0) {
                    const int pb = (m - 1) >> 2;
                    const float tps = pb == 0 ? tb[0] : pb == 1 ? tb[1] : pb == 2 ? tb[2] : tb[3];
                    const float fps = pb == 0 ? fv[0] : pb == 1 ? fv[1] : pb == 2 ? fv[2] : fv[3];
                    t_prev = __shfl(tps, gbase + ((m - 1) & 3), 64);
                    f_t = __shfl(fps, gbase + ((m - 1) & 3), 64);
                    have_prev = true;
                }
                t = tm;
                if (!(tm < tfar)) {  // the march ends before sample m
                    n_eval += m;
                    break;
                }
                f_tt = fm;  // < 0: got it
                n_eval += m + 1;
                break;
            }
            if (f_tt < 0.0f) {
                if (!have_prev) f_t = interp_vol<S, false, DP>(V, a, o[0] + d[0] * t_prev, o[1] + d[1] * t_prev, o[2] + d[2] * t_prev);
                t = t + stepsize * f_tt / (f_t - f_tt);
                hp[0] = o[0] + d[0] * t;
                hp[1] = o[1] + d[1] * t;
                hp[2] = o[2] + d[2] * t;
                th = t;
                hit = true;
            }
        }
    }
    const unsigned cost = n_brick + (n_eval + 3u) / 4u;
    if (q != 0) return cost;
    const int i = y * a.w + x;
    if (hit && th > 0.0f) {
        a.refV[i] = make_float4(hp[0], hp[1], hp[2], 1.0f);
        float g[3];
        gradient(V, a, hp, g);
        const float len = sqrtf(dot3(g[0], g[1], g[2], g[0], g[1], g[2]));
        if (len == 0.0f)
            a.refN[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        else
            a.refN[i] = make_float4(g[0] / len, g[1] / len, g[2] / len, 1.0f);
    } else {
        a.refV[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.refN[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (a.stats) {
        atomicAdd(&a.stats[0], (unsigned long long)n_eval);
        atomicAdd(&a.stats[1], (unsigned long long)n_skip);
        atomicAdd(&a.stats[2], (unsigned long long)n_brick);
        atomicAdd(&a.stats[3], (unsigned long long)(hit && th > 0.0f));
        atomicMax(&a.stats[4], (unsigned long long)n_eval);
        atomicMax(&a.stats[5], (unsigned long long)(n_eval + n_brick));
    }
    return cost;
}

// Four lanes per ray over 8x8-pixel tiles, 64 rays per block
template <class S>
__device__ __forceinline__ void raycast_coop_body(const RaycastArgs &a, const DevState *__restrict__ st, int blk)
{
    const int nb = a.tiles_x * a.tiles_y;
    const int G = a.xcd_group;
    const int j = blk / 8;
    const int tile = ((blk % 8) + 8 * (j / G)) * G + (j % G);
    if (tile >= nb) return;
    const int q = threadIdx.x & 3, r = threadIdx.x >> 2;
    const int gbase = (int)(threadIdx.x & 63) & ~3;  // the group's first lane in the wavefront
    const int x = (tile % a.tiles_x) * 8 + (r & 7);
    const int y = (tile / a.tiles_x) * 8 + (r >> 3);
    if (x >= a.w || y >= a.h) return;  // the whole group
    float view[16];
    raycast_view(a, st, view);
    (void)raycast_coop_ray<S>(a, view, x, y, q, gbase);
}



// Diagnostic: the record of this wavefront in a.wave_trace (8 words at
// 4 x blk + wave: start and end clock, HW_ID, XCC id, tile, and 0xAB << 24 |
// lanes per ray << 16 | the wavefront's largest loop iteration count, capped
// at 65535).  The first active lane writes the first seven words with vector
// stores; every active lane folds its count into the last word (atomicMax
// over words with the same top 16 bits, zeroed before the launch).
__device__ __forceinline__ void wave_trace_record(const RaycastArgs &a, int blk, unsigned long long wt0, int tile,
                                                  int lanes, unsigned cost)
{
    const unsigned long long wt1 = __builtin_amdgcn_s_memrealtime();
    unsigned hw, xcc;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_HW_ID)" : "=s"(hw));
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
    uint32_t *rec = a.wave_trace + ((size_t)blk * 4 + (threadIdx.x >> 6)) * 8;
    const unsigned long long act = __ballot(1);
    if ((int)(threadIdx.x & 63) == __ffsll((long long)act) - 1) {
        rec[0] = (uint32_t)wt0;
        rec[1] = (uint32_t)(wt0 >> 32);
        rec[2] = (uint32_t)wt1;
        rec[3] = (uint32_t)(wt1 >> 32);
        rec[4] = hw;
        rec[5] = xcc;
        rec[6] = (uint32_t)tile;
    }
    atomicMax(&rec[7], 0xAB000000u | ((uint32_t)lanes << 16) | min(cost, 0xFFFFu));
}

// One thread per pixel; 16x16 pixel tiles per block.  Block ids are remapped
// so that the blocks an XCD receives (id mod 8) cover one contiguous band of
// tiles, keeping the voxels neighbouring rays share in that XCD's L2.
template <class S, bool TOUCH = false, int DP = -1>
__device__ __forceinline__ void raycast_body(const RaycastArgs &a, const DevState *__restrict__ st, int blk)
{
    // blocks are dealt to the 8 XCDs round-robin (from a starting XCD that
    // varies between launches), so blocks whose numbers agree mod 8 share an
    // XCD; residue x receives the clusters of xcd_group consecutive tiles x,
    // x + 8, x + 16, ... (1 = fully interleaved)
    const int nb = a.tiles_x * a.tiles_y;
    const int G = a.xcd_group;
    const int j = blk / 8;
    const int tile = ((blk % 8) + 8 * (j / G)) * G + (j % G);
    if (tile >= nb) return;
    const int x = (tile % a.tiles_x) * 16 + (threadIdx.x & 15);
    const int y = (tile / a.tiles_x) * 16 + (threadIdx.x >> 4);
    if (x >= a.w || y >= a.h) return;
    const unsigned long long wt0 = a.wave_trace ? __builtin_amdgcn_s_memrealtime() : 0ull;
    float view[16];
    raycast_view(a, st, view);
    const unsigned cost = raycast_pixel<S, TOUCH, DP>(a, view, x, y);
    if (a.wave_trace) wave_trace_record(a, blk, wt0, tile, 1, cost);
}

template <class S, int DP>
__global__ __launch_bounds__(256) void k_raycast(RaycastArgs a, const DevState *__restrict__ st)
{
    raycast_body<S, false, DP>(a, st, blockIdx.x);
}

// Hybrid schedule.  The raycast's time is set by its slowest wavefronts,
// which hold grazing rays and cluster in a few tiles; their fine-step
// samples spread over four lanes per ray (raycast_coop_ray, the same bits)
// cut those tiles' chains, while every other tile keeps one lane per ray.
// Which tiles: those whose rays took the most loop iterations in the previous
// launch of this session (consecutive frames see nearly the same scene).  The
// grid is 4 x heavy_max blocks for the quarters of the heavy tiles (the four
// quarters of a tile get block numbers equal mod 8, so they share an XCD),
// dispatched first, then the one-lane blocks, which skip heavy tiles.  Every
// wavefront records its tile's largest one-lane iteration count; the
// session's next pyramid launch picks the next launch's heavy tiles from them
// (select_heavy_tiles, beside the pyramid's blocks, off the raycast's
// critical path).
template <class S, int DP>
__global__ __launch_bounds__(256) void k_raycast_hybrid(RaycastArgs a, const DevState *__restrict__ st)
{
    const int nheavy_blocks = 4 * a.heavy_max;
    const int b = blockIdx.x;
    const unsigned long long wt0 = a.wave_trace ? __builtin_amdgcn_s_memrealtime() : 0ull;
    unsigned cost = 0;
    bool in_img = false;  // this lane's ray is a pixel of the image
    int tile = -1;  // the 16x16 tile this wavefront's rays belong to
    float view[16];
    if (b >= nheavy_blocks || (b / 32) * 8 + (b % 8) < min(a.heavy[0], a.heavy_max)) raycast_view(a, st, view);
    if (b < nheavy_blocks) {
        // quarter qd of heavy tile h, 64 rays of four lanes
        const int h = (b / 32) * 8 + (b % 8), qd = (b / 8) % 4;
        if (h < min(a.heavy[0], a.heavy_max)) {
            tile = a.heavy[1 + h];
            const int q = threadIdx.x & 3, r = threadIdx.x >> 2;
            const int gbase = (int)(threadIdx.x & 63) & ~3;
            const int x = (tile % a.tiles_x) * 16 + (qd & 1) * 8 + (r & 7);
            const int y = (tile / a.tiles_x) * 16 + (qd >> 1) * 8 + (r >> 3);
            in_img = x < a.w && y < a.h;
            if (in_img) cost = raycast_coop_ray<S, DP>(a, view, x, y, q, gbase);
        }
    } else {
        const int blk = b - nheavy_blocks;
        const int nb = a.tiles_x * a.tiles_y;
        const int G = a.xcd_group;
        const int j = blk / 8;
        const int t = ((blk % 8) + 8 * (j / G)) * G + (j % G);
        if (t < nb && !a.heavy_flag[t]) {
            tile = t;
            const int x = (t % a.tiles_x) * 16 + (threadIdx.x & 15);
            const int y = (t / a.tiles_x) * 16 + (threadIdx.x >> 4);
            in_img = x < a.w && y < a.h;
            if (in_img) cost = raycast_pixel<S, false, DP>(a, view, x, y);
        }
    }
    // the wavefront's largest count into its tile's (tile is uniform per wavefront)
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) cost = max(cost, (unsigned)__shfl_xor((int)cost, o, 64));
    if ((threadIdx.x & 63) == 0 && tile >= 0 && cost > 0u) atomicMax(&a.tile_cost[tile], cost);
    if (a.wave_trace && tile >= 0 && __any(in_img)) wave_trace_record(a, b, wt0, tile, b < nheavy_blocks ? 4 : 1, cost);
}

// k_raycast with every voxel and flag byte read marked in a.touch (counts the
// distinct bytes the march reads; results identical to k_raycast)
template <class S>
__global__ __launch_bounds__(256) void k_raycast_touch(RaycastArgs a, const DevState *__restrict__ st)
{
    raycast_body<S, true>(a, st, blockIdx.x);
}

// population count of n words, added to *out
__global__ __launch_bounds__(256) void k_popcount(const uint32_t *__restrict__ w, size_t n,
                                                  unsigned long long *__restrict__ out)
{
    unsigned c = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        c += __popc(w[i]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, (unsigned long long)c);
}

template <class S>
__global__ __launch_bounds__(256) void k_raycast_coop(RaycastArgs a, const DevState *__restrict__ st)
{
    raycast_coop_body<S>(a, st, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Batched kernels: block -> (session, local block).  Sessions of a batch share
// the image size, ICP schedule, storage type and arithmetic mode; volume sizes
// may differ (integrate / brick launches map blocks through prefix tables).
// Blocks of sessions without a frame in this launch return at once.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool batch_active(const BatchTick *tp, int i) { return (tp->active >> i) & 1ull; }
// the tick record (depth pointers, active mask) is read from device memory, so
// that a captured launch sequence can be replayed for every tick; the member
// tables are kernel arguments

// largest i with prefix[i] <= b
__device__ __forceinline__ int batch_find(const int *prefix, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_batch_pyramid(BatchTable tb, const BatchTick *__restrict__ tp, int per)
{
    const int i = blockIdx.x / per;
    if (!batch_active(tp, i)) return;
    // by reference: the body indexes the level arrays with a run-time level,
    // which a private copy of the arguments would put in scratch memory
    pyramid_body<uint16_t>(tb.sess[i].pyr, tp->depth[i], blockIdx.x % per);
}

__global__ __launch_bounds__(kIcpThreads) void k_batch_icp(BatchTable tb, const BatchTick *__restrict__ tp, int level,
                                                           int first, int save_pose, int finish, int nblk)
{
    const int i = blockIdx.x / nblk;
    if (!batch_active(tp, i)) return;
    IcpArgs a = tb.sess[i].icp[level];
    a.first = first;
    a.save_pose = save_pose;
    a.finish = finish;
    a.frame = 1;
    icp_body(a, tb.sess[i].st, blockIdx.x % nblk, nblk);
}

// Track() of every active session in one persistent launch: session i owns
// blocks [i * per, (i + 1) * per) and runs icp_track_body over them (its own
// state, rows and tags), so its result is that of its launch alone.  The
// whole grid fits the device at once (per * n <= compute units).
__global__ __launch_bounds__(kIcpThreads) void k_batch_icp_track(BatchTable tb, const BatchTick *__restrict__ tp,
                                                                 int per)
{
    const int i = blockIdx.x / per;
    if (!batch_active(tp, i)) return;
    icp_track_body(tb.sess[i].trk, tb.sess[i].st, blockIdx.x % per, per);
}

template <class S, bool FAST, bool OFF32>
__global__ __launch_bounds__(256) void k_batch_integrate(BatchTable tb, const BatchTick *__restrict__ tp)
{
    const int i = batch_find(tb.int_prefix, tb.n, blockIdx.x);
    if (!batch_active(tp, i)) return;
    const IntegrateArgs a = tb.sess[i].integ;
    const int gx = tb.sess[i].int_grid[0], gy = tb.sess[i].int_grid[1];
    const int local = blockIdx.x - tb.int_prefix[i];
    const int bx = local % gx, by = (local / gx) % gy, bz = local / (gx * gy);
    if (FAST) integrate_fast_body<S, OFF32>(a, tb.sess[i].st, bx, by, bz);
    else integrate_body<S, false>(a, tb.sess[i].st, bx, by, bz);
}

__global__ __launch_bounds__(256) void k_batch_dilate(BatchTable tb, const BatchTick *__restrict__ tp)
{
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);  // one wavefront per coarse cell
    if (b >= tb.coa_prefix[tb.n]) return;
    const int i = batch_find(tb.coa_prefix, tb.n, b);
    if (!batch_active(tp, i)) return;
    const BrickArgs a = tb.sess[i].brick;
    brick_dilate_body(a, b - tb.coa_prefix[i]);
}

// blocks_per_session is a multiple of 8, so a block's global and
// session-local numbers agree mod 8 (the residue that decides its XCD) and
// the per-session tile clustering holds
template <class S, int DP>
__global__ __launch_bounds__(256) void k_batch_raycast(BatchTable tb, const BatchTick *__restrict__ tp, int per)
{
    const int i = blockIdx.x / per;
    if (!batch_active(tp, i)) return;
    const RaycastArgs a = tb.sess[i].ray;
    raycast_body<S, false, DP>(a, tb.sess[i].st, blockIdx.x % per);
}

template <class S>
__global__ __launch_bounds__(256) void k_batch_raycast_coop(BatchTable tb, const BatchTick *__restrict__ tp, int per)
{
    const int i = blockIdx.x / per;
    if (!batch_active(tp, i)) return;
    const RaycastArgs a = tb.sess[i].ray;
    raycast_coop_body<S>(a, tb.sess[i].st, blockIdx.x % per);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// Plain launch, or with start / stop events carried by the dispatch packet.
template <typename... KArgs, typename... Args>
static void launch_timed(void (*k)(KArgs...), dim3 g, dim3 b, hipStream_t s, hipEvent_t t0, hipEvent_t t1,
                         Args... args)
{
    if (t0 || t1) hipExtLaunchKernelGGL(k, g, b, 0, s, t0, t1, 0, args...);
    else hipLaunchKernelGGL(k, g, b, 0, s, args...);
}

hipError_t launch_pyramid_u16(const PyramidArgs &a, const uint16_t *in, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    launch_timed(k_pyramid<uint16_t>, dim3(a.blocks_total), dim3(256), s, t0, t1, a, in);
    return hipGetLastError();
}
hipError_t launch_pyramid_f32(const PyramidArgs &a, const float *in, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    launch_timed(k_pyramid<float>, dim3(a.blocks_total), dim3(256), s, t0, t1, a, in);
    return hipGetLastError();
}
hipError_t launch_icp_iteration(const IcpArgs &a, DevState *st, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    // 8 row groups x rows_step blocks; each block takes every rows_step-th row
    // of its group (fewer, fuller blocks: one release + one ticket each)
    launch_timed(k_icp_iteration, dim3(icp_blocks(a)), dim3(kIcpThreads), s, t0, t1, a, st);
    return hipGetLastError();
}
hipError_t launch_icp_track(const IcpTrackArgs &a, DevState *st, int nblk, hipStream_t s, hipEvent_t t0,
                            hipEvent_t t1)
{
    for (int l = 0; l < 3; ++l)
        if (a.iters[l] > 0 && icp_blocks(a.lv[l]) > nblk) return hipErrorInvalidValue;
    launch_timed(k_icp_track, dim3(nblk), dim3(kIcpThreads), s, t0, t1, a, st);
    return hipGetLastError();
}

hipError_t launch_icp_ref_iteration(const IcpArgs &a, DevState *st, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (!a.dbg_result || !a.dbg_error || !a.dbg_J) return hipErrorInvalidValue;
    launch_timed(k_icp_records, dim3((a.w * a.h + 255) / 256), dim3(256), s, t0, nullptr, a, (const DevState *)st);
    launch_timed(k_icp_reduce_ref, dim3(8), dim3(128), s, nullptr, t1, a, st);
    return hipGetLastError();
}

hipError_t launch_icp_finish(DevState *st, int frame, float *lib, int lib_max, hipStream_t s, hipEvent_t t1)
{
    launch_timed(k_icp_finish, dim3(1), dim3(64), s, nullptr, t1, st, frame, lib, lib_max);
    return hipGetLastError();
}
hipError_t launch_state_set(DevState *st, const float *pose, int flags, hipStream_t s)
{
    Mat16 m;
    for (int i = 0; i < 16; ++i) m.m[i] = pose ? pose[i] : 0.0f;
    if (pose) flags |= kStateSetPose;
    else flags &= ~(kStateSetPose | kStateResetPose);
    hipLaunchKernelGGL(k_state_set, dim3(1), dim3(64), 0, s, st, m, flags);
    return hipGetLastError();
}

static dim3 integrate_grid(const IntegrateArgs &a)
{
    return dim3((a.N[0] + 63) / 64, (a.N[1] + a.rows - 1) / a.rows, (a.N[2] + a.zlen - 1) / a.zlen);
}

void integrate_grid_dims(const IntegrateArgs &a, int g[3])
{
    const dim3 d = integrate_grid(a);
    g[0] = (int)d.x;
    g[1] = (int)d.y;
    g[2] = (int)d.z;
}

hipError_t launch_batch_pyramid(const BatchTable &tb, const BatchTick *tp, int per, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_pyramid, dim3(tb.n * per), dim3(256), 0, s, tb, tp, per);
    return hipGetLastError();
}

hipError_t launch_batch_icp(const BatchTable &tb, const BatchTick *tp, int level, int first, int save_pose,
                            int finish, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_icp, dim3(tb.n * nblk), dim3(kIcpThreads), 0, s, tb, tp, level, first, save_pose,
                       finish, nblk);
    return hipGetLastError();
}

hipError_t launch_batch_icp_track(const BatchTable &tb, const BatchTick *tp, int per, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_icp_track, dim3(tb.n * per), dim3(kIcpThreads), 0, s, tb, tp, per);
    return hipGetLastError();
}

hipError_t launch_batch_integrate(const BatchTable &tb, const BatchTick *tp, int storage, int fast, int off32,
                                  int total, int rows, hipStream_t s)
{
    const dim3 g(total), b(64, rows);
#define KF_BINT(ST)                                                                                  \
    if (fast && off32) hipLaunchKernelGGL((k_batch_integrate<ST, true, true>), g, b, 0, s, tb, tp);  \
    else if (fast) hipLaunchKernelGGL((k_batch_integrate<ST, true, false>), g, b, 0, s, tb, tp);     \
    else hipLaunchKernelGGL((k_batch_integrate<ST, false, false>), g, b, 0, s, tb, tp);
    switch (storage) {
    case KF_STORAGE_RG16I: KF_BINT(StoreRG16I); break;
    case KF_STORAGE_F32: KF_BINT(StoreF32); break;
    case KF_STORAGE_F16: KF_BINT(StoreF16); break;
    default: return hipErrorInvalidValue;
    }
#undef KF_BINT
    return hipGetLastError();
}

hipError_t launch_batch_bricks(const BatchTable &tb, const BatchTick *tp, int total_coarse, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_dilate, dim3((total_coarse + 3) / 4), dim3(256), 0, s, tb, tp);
    return hipGetLastError();
}

hipError_t launch_batch_raycast(const BatchTable &tb, const BatchTick *tp, int storage, int lanes, int per,
                                int dim_pow2, hipStream_t s)
{
    const dim3 g(tb.n * per), b(256);
    if (lanes == 4) {
        switch (storage) {
        case KF_STORAGE_RG16I: hipLaunchKernelGGL(k_batch_raycast_coop<StoreRG16I>, g, b, 0, s, tb, tp, per); break;
        case KF_STORAGE_F32: hipLaunchKernelGGL(k_batch_raycast_coop<StoreF32>, g, b, 0, s, tb, tp, per); break;
        case KF_STORAGE_F16: hipLaunchKernelGGL(k_batch_raycast_coop<StoreF16>, g, b, 0, s, tb, tp, per); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    // dim_pow2: every member's dims are powers of two (else the kernel
    // reads each member's flag)
#define KF_BRC(ST)                                                                            \
    if (dim_pow2) hipLaunchKernelGGL((k_batch_raycast<ST, 1>), g, b, 0, s, tb, tp, per);     \
    else hipLaunchKernelGGL((k_batch_raycast<ST, -1>), g, b, 0, s, tb, tp, per);
    switch (storage) {
    case KF_STORAGE_RG16I: KF_BRC(StoreRG16I); break;
    case KF_STORAGE_F32: KF_BRC(StoreF32); break;
    case KF_STORAGE_F16: KF_BRC(StoreF16); break;
    default: return hipErrorInvalidValue;
    }
#undef KF_BRC
    return hipGetLastError();
}

hipError_t launch_integrate(IntegrateArgs a, int storage, int fast, const DevState *st, hipStream_t s,
                            hipEvent_t t0, hipEvent_t t1)
{
    const dim3 g = integrate_grid(a), b(64, a.rows);
    const bool off32 = integrate_planes_fit_32(a, storage);
    if (fast && a.col_pos) {
        // the timed interval covers the column set-up and the integrate
        launch_timed(k_integrate_columns, dim3((a.N[0] + 63) / 64, (a.N[1] + 3) / 4), dim3(64, 4), s, t0,
                     (hipEvent_t) nullptr, a, st);
        t0 = nullptr;
    } else {
        a.col_pos = nullptr;
    }
#define KF_INT(ST)                                                                     \
    if (fast && off32) launch_timed(k_integrate_fast<ST, true>, g, b, s, t0, t1, a, st);  \
    else if (fast) launch_timed(k_integrate_fast<ST, false>, g, b, s, t0, t1, a, st);     \
    else launch_timed(k_integrate<ST, false>, g, b, s, t0, t1, a, st);
    switch (storage) {
    case KF_STORAGE_RG16I: KF_INT(StoreRG16I); break;
    case KF_STORAGE_F32: KF_INT(StoreF32); break;
    case KF_STORAGE_F16: KF_INT(StoreF16); break;
    default: return hipErrorInvalidValue;
    }
#undef KF_INT
    return hipGetLastError();
}

hipError_t launch_brick_rebuild(const BrickArgs &a, int storage, hipStream_t s)
{
    const dim3 g((a.N[0] + 63) / 64, a.NB[1], a.NB[2]), b(512);
    switch (storage) {
    case KF_STORAGE_RG16I: hipLaunchKernelGGL(k_brick_rebuild<StoreRG16I>, g, b, 0, s, a); break;
    case KF_STORAGE_F32: hipLaunchKernelGGL(k_brick_rebuild<StoreF32>, g, b, 0, s, a); break;
    case KF_STORAGE_F16: hipLaunchKernelGGL(k_brick_rebuild<StoreF16>, g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_brick_dilate(const BrickArgs &a, hipStream_t s)
{
    const int nc = a.NC[0] * a.NC[1] * a.NC[2];
    hipLaunchKernelGGL(k_brick_dilate, dim3((nc + 3) / 4), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_count_integrate_fast(const IntegrateArgs &a, int storage, const DevState *st,
                                       unsigned long long *counts, hipStream_t s)
{
    const dim3 g = integrate_grid(a), b(64, a.rows);
    const bool off32 = integrate_planes_fit_32(a, storage);
#define KF_INT(ST)                                                                               \
    if (off32) hipLaunchKernelGGL((k_count_integrate_fast<ST, true>), g, b, 0, s, a, st, counts); \
    else hipLaunchKernelGGL((k_count_integrate_fast<ST, false>), g, b, 0, s, a, st, counts);
    switch (storage) {
    case KF_STORAGE_RG16I: KF_INT(StoreRG16I); break;
    case KF_STORAGE_F32: KF_INT(StoreF32); break;
    case KF_STORAGE_F16: KF_INT(StoreF16); break;
    default: return hipErrorInvalidValue;
    }
#undef KF_INT
    return hipGetLastError();
}

hipError_t launch_count_integrate(const IntegrateArgs &a, const DevState *st, unsigned long long *counts,
                                  hipStream_t s)
{
    const dim3 g((a.N[0] + 63) / 64, (a.N[1] + 3) / 4, 1), b(64, 4);
    hipLaunchKernelGGL(k_count_integrate, g, b, 0, s, a, st, counts);
    return hipGetLastError();
}

// intensityProjection (gFusion.cpp:1368-1409, mip.cs:33-107): per pixel, the
// sum of the positive trilinear samples of the TSDF along the ray at t =
// tnear, tnear + largeStep, ... < tfar; written as (sum, 0, 0, 1).
//
// kMipLanes lanes share a ray.  Every lane of the group steps t through the
// same float additions; lane l evaluates the samples j = l (mod kMipLanes),
// kMipBatch of them per round with all their loads issued before the first
// is used, and every lane of the group then adds the round's positive values
// in sample order (broadcast within the group), so the sum is the sequential
// one, bit for bit.  With a.skip, a.dflags holds the dilated zero flags: a
// sample whose lower corner lies in a brick with a clear flag reads only
// (0, 0) voxels, its value is 0, it adds nothing and is not evaluated.
constexpr int kMipLanes = 4, kMipBatch = 2;

template <class S>
__global__ __launch_bounds__(256) void k_intensity_projection(RaycastArgs a, const DevState *__restrict__ st,
                                                              float4 *__restrict__ out)
{
    constexpr int L = kMipLanes, B = kMipBatch, R = 256 / L;  // R rays: 8 x (R / 8) pixels
    const int lane = threadIdx.x & (L - 1), ray = threadIdx.x / L;
    const int x = blockIdx.x * 8 + (ray & 7), y = blockIdx.y * (R / 8) + (ray >> 3);
    if (x >= a.w || y >= a.h) return;  // the whole group of L lanes
    float view[16];
    {
        float iK[16], P[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) P[i] = st->pose[i];
        inverse_camera_matrix(a.cam, iK);
        mul4(P, iK, view);  // m_pose * invK (gFusion.cpp:1376)
    }
    const VolView<S> V(a);
    const float o[3] = {view[12], view[13], view[14]};
    float d[3];
    rotate3(view, (float)x, (float)y, 1.0f, d);
    float tmin[3], tmax[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float invR = 1.0f / d[k];
        const float tbot = (-1.0f * invR) * o[k];
        const float ttop = invR * (a.dim[k] - o[k]);
        tmin[k] = gmin(ttop, tbot);
        tmax[k] = gmax(ttop, tbot);
    }
    const float largest_tmin = gmax(gmax(tmin[0], tmin[1]), gmax(tmin[0], tmin[2]));
    const float smallest_tmax = gmin(gmin(tmax[0], tmax[1]), gmin(tmax[0], tmax[2]));
    const float tnear = gmax(largest_tmin, a.near_plane);
    const float tfar = gmin(smallest_tmax, a.far_plane);
    float sum = 0.0f;
    // tnear, tfar and t are the same in every lane of the group, so the
    // group enters, loops and leaves together
    float t = tnear;
    int guard = 0;
    while (t < tfar && guard < kMaxMarchSteps) {
        // this lane's samples of the round: j = b L + lane
        float tj[B];
        bool ok[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            tj[b] = 0.0f;
            ok[b] = false;
#pragma unroll
            for (int l = 0; l < L; ++l) {
                if (l == lane) {
                    tj[b] = t;
                    ok[b] = t < tfar && guard < kMaxMarchSteps;
                }
                t += a.large_step;
                ++guard;
            }
        }
        FastSample q[B];
        bool ev[B], fast = a.off32 != 0;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            sample_fast(a, o[0] + d[0] * tj[b], o[1] + d[1] * tj[b], o[2] + d[2] * tj[b], V.plane32, q[b]);
            ev[b] = ok[b];
            if (a.skip && ok[b] && q[b].inside)
                ev[b] = a.dflags[((q[b].b[2] >> 3) * a.NB[1] + (q[b].b[1] >> 3)) * a.NB[0] + (q[b].b[0] >> 3)] != 0;
            fast = fast && (!ev[b] || q[b].inside);
        }
        float fv[B];
        if (__all(fast)) {
            // corners b and b + 1 on every axis: four x-pairs per sample,
            // from the four uniform row bases
            constexpr uint32_t sz = VolView<S>::kSz;
            const size_t rowb = (size_t)V.N0 * sz, plb = (size_t)V.plane32 * sz;
            const char *rb0 = (const char *)V.v;
            const char *rb1 = rb0 + rowb, *rb2 = rb0 + plb, *rb3 = rb0 + (plb + rowb);
            float c[B][8];
#pragma unroll
            for (int b = 0; b < B; ++b) {
#pragma unroll
                for (int k = 0; k < 8; ++k) c[b][k] = 0.0f;
                if (ev[b]) {
                    const uint32_t vo = q[b].base * sz;
                    V.pair(rb0, vo, c[b][0], c[b][1]);
                    V.pair(rb1, vo, c[b][2], c[b][3]);
                    V.pair(rb2, vo, c[b][4], c[b][5]);
                    V.pair(rb3, vo, c[b][6], c[b][7]);
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b)
                fv[b] = ev[b] ? trilerp<S>(c[b][0], c[b][1], c[b][2], c[b][3], c[b][4], c[b][5], c[b][6], c[b][7],
                                           q[b].f)
                              : 0.0f;
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                fv[b] = 0.0f;
                if (ev[b]) fv[b] = interp_vol(V, a, o[0] + d[0] * tj[b], o[1] + d[1] * tj[b], o[2] + d[2] * tj[b]);
            }
        }
        // the round's samples in order: j = b L + l is lane l's sample b
#pragma unroll
        for (int b = 0; b < B; ++b) {
#pragma unroll
            for (int l = 0; l < L; ++l) {
                const float f = __shfl(fv[b], l, L);
                if (f > 0.0f) sum += f;
            }
        }
    }
    if (lane == 0) out[(size_t)y * a.w + x] = make_float4(sum, 0.0f, 0.0f, 1.0f);
}

hipError_t launch_intensity_projection(const RaycastArgs &a, int storage, const DevState *st, float4 *out,
                                       hipStream_t s)
{
    const dim3 g((a.w + 7) / 8, (a.h + 256 / kMipLanes / 8 - 1) / (256 / kMipLanes / 8)), b(256);
    switch (storage) {
    case KF_STORAGE_RG16I: hipLaunchKernelGGL(k_intensity_projection<StoreRG16I>, g, b, 0, s, a, st, out); break;
    case KF_STORAGE_F32: hipLaunchKernelGGL(k_intensity_projection<StoreF32>, g, b, 0, s, a, st, out); break;
    case KF_STORAGE_F16: hipLaunchKernelGGL(k_intensity_projection<StoreF16>, g, b, 0, s, a, st, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_raycast_touch(const RaycastArgs &a, int storage, const DevState *st, size_t words,
                                unsigned long long *count, hipStream_t s)
{
    const dim3 g(a.blocks), b(256);
    switch (storage) {
    case KF_STORAGE_RG16I: hipLaunchKernelGGL(k_raycast_touch<StoreRG16I>, g, b, 0, s, a, st); break;
    case KF_STORAGE_F32: hipLaunchKernelGGL(k_raycast_touch<StoreF32>, g, b, 0, s, a, st); break;
    case KF_STORAGE_F16: hipLaunchKernelGGL(k_raycast_touch<StoreF16>, g, b, 0, s, a, st); break;
    default: return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL(k_popcount, dim3(1024), dim3(256), 0, s, a.touch, words, count);
    return hipGetLastError();
}

hipError_t launch_raycast(const RaycastArgs &a, int storage, const DevState *st, hipStream_t s, hipEvent_t t0,
                          hipEvent_t t1)
{
    const dim3 g(a.blocks), b(256);
    if (a.lanes == 4) {
        switch (storage) {
        case KF_STORAGE_RG16I: launch_timed(k_raycast_coop<StoreRG16I>, g, b, s, t0, t1, a, st); break;
        case KF_STORAGE_F32: launch_timed(k_raycast_coop<StoreF32>, g, b, s, t0, t1, a, st); break;
        case KF_STORAGE_F16: launch_timed(k_raycast_coop<StoreF16>, g, b, s, t0, t1, a, st); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    // the hybrid schedule, unless a diagnostic needs the one-lane kernel's counters
    const bool hybrid = a.heavy_max > 0 && a.tile_cost && !a.stats && !a.ray_counts;
    const dim3 gh(a.blocks + 4 * a.heavy_max);
#define KF_RC(ST)                                                                  \
    if (hybrid && a.dim_pow2) launch_timed(k_raycast_hybrid<ST, 1>, gh, b, s, t0, t1, a, st); \
    else if (hybrid) launch_timed(k_raycast_hybrid<ST, 0>, gh, b, s, t0, t1, a, st);         \
    else if (a.dim_pow2) launch_timed(k_raycast<ST, 1>, g, b, s, t0, t1, a, st);              \
    else launch_timed(k_raycast<ST, 0>, g, b, s, t0, t1, a, st);
    switch (storage) {
    case KF_STORAGE_RG16I: KF_RC(StoreRG16I); break;
    case KF_STORAGE_F32: KF_RC(StoreF32); break;
    case KF_STORAGE_F16: KF_RC(StoreF16); break;
    default: return hipErrorInvalidValue;
    }
#undef KF_RC
    return hipGetLastError();
}

}  // namespace kf
