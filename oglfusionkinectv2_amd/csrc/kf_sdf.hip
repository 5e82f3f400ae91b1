// kf_sdf.hip -- point-to-volume tracking (TrackSDF) and pose recovery for gfx950.
//
//   k_sdf_track  one persistent launch runs a whole TrackSDF() (gFusion.cpp:
//                835-1038: trackSDF.cs + reduceSDF.cs + getSDFReduction + the
//                scaled, damped LDL^T step, every iteration of every level) or
//                a whole recoverPose() (gFusion.cpp:1616-1653: TrackSDF from
//                each pose-library entry until one succeeds).
//
// The grid is at most one block per CU, so every block is resident at once;
// the blocks meet at one grid barrier per iteration (an arrival counter in the
// session state, agent-scope release / acquire).  After the barrier every
// block sums the block partials itself in the same fixed order and takes the
// same Gauss-Newton step, so all blocks agree on the pose, the convergence
// test and hence on the number of barriers without a second barrier.  Each
// barrier wait is bounded: a block that waits too long raises an abort flag
// that ends every block at its next barrier, and the session reports it.
//
// trackSDF.cs writes its 32-byte record only for x < width - 1 (:455); the
// records of the last column keep whatever an earlier dispatch at any level
// left at that buffer index, and reduceSDF.cs sums them.  Those indices are
// y * W_l + W_l - 1 for the three levels (at most H_0 + H_1 + H_2 of them);
// the kernel keeps one record slot per such index and writes a computed
// record to its slot whenever the pixel's index is one of them, so the stale
// records read are exactly the reference's without a full-image buffer.
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "kf_device.hpp"
#include "kf_dilate.hpp"
#include "kf_internal.hpp"
#include "kf_store.hpp"

namespace kf {

constexpr int kSdfThreads = 256;
constexpr int kSdfTerms = 28;  // 2 D^2, h J D [6], h J J^T upper triangle [21]

// GLSL modf: integer part toward zero and a fraction with the sign of x
__device__ __forceinline__ float gmodf(float x, float &ip)
{
    ip = truncf(x);
    if (isinf(x)) return copysignf(0.0f, x);
    return copysignf(x - ip, x);
}

// ivec() of a whole-number float; NaN / out of range give INT_MIN, whose
// image loads are out of bounds and read zero
__device__ __forceinline__ int sdf_f2i(float v)
{
    if (!(v > -2147483648.0f && v < 2147483648.0f)) return INT32_MIN;
    return (int)v;
}

__device__ __forceinline__ int sdf_ip1(int v) { return v == INT32_MIN ? v : v + 1; }

template <class S>
__device__ __forceinline__ float sdf_vox(const typename S::V *__restrict__ vol, const SdfArgs &a, int x, int y, int z)
{
    if ((unsigned)x >= (unsigned)a.N[0] || (unsigned)y >= (unsigned)a.N[1] || (unsigned)z >= (unsigned)a.N[2])
        return 0.0f;
    return S::value(vol[((size_t)z * a.N[1] + y) * a.N[0] + x]);
}

// SDF(location), trackSDF.cs:186-220
template <class S>
__device__ __forceinline__ float sdf_sample(const typename S::V *__restrict__ vol, const SdfArgs &a, float lx, float ly,
                                            float lz)
{
    float i, j, k;
    const float x = gmodf(lx, i), y = gmodf(ly, j), z = gmodf(lz, k);
    const int I = sdf_f2i(i), J = sdf_f2i(j), K = sdf_f2i(k);
    const int I1 = sdf_ip1(I), J1 = sdf_ip1(J), K1 = sdf_ip1(K);
    const float N0 = sdf_vox<S>(vol, a, I, J, K);
    const float N1 = sdf_vox<S>(vol, a, I, J1, K);
    const float N2 = sdf_vox<S>(vol, a, I1, J, K);
    const float N3 = sdf_vox<S>(vol, a, I1, J1, K);
    const float N4 = sdf_vox<S>(vol, a, I, J, K1);
    const float N5 = sdf_vox<S>(vol, a, I, J1, K1);
    const float N6 = sdf_vox<S>(vol, a, I1, J, K1);
    const float N7 = sdf_vox<S>(vol, a, I1, J1, K1);
    const float a0 = N0 * (1.0f - z) + N4 * z;
    const float a1 = N1 * (1.0f - z) + N5 * z;
    const float b0 = N2 * (1.0f - z) + N6 * z;
    const float b1 = N3 * (1.0f - z) + N7 * z;
    return (a0 * (1.0f - y) + a1 * y) * (1.0f - x) + (b0 * (1.0f - y) + b1 * y) * x;
}

// trackSDF.cs:460-530 for one pixel: rec = {h, D, J[6]}
template <class S>
__device__ __forceinline__ void sdf_record(const typename S::V *__restrict__ vol, const SdfArgs &a, const float *T,
                                           float4 v, float rec[8])
{
    const float cp[4] = {v.x, v.y, v.z, 1.0f};
    float pv[4];
    matvec4(T, cp, pv);
    const float c0 = ((pv[0] * (float)a.N[0] / a.dim[0]) - 0.5f) - 0.5f;
    const float c1 = ((pv[1] * (float)a.N[1] / a.dim[1]) - 0.5f) - 0.5f;
    const float c2 = (pv[2] * (float)a.N[2] / a.dim[2]) - 0.5f;
    const float D = sdf_sample<S>(vol, a, c0, c1, c2);
    const float den = 2.0f * (a.dim[0] / (float)a.N[0]);
    const float gx = (sdf_sample<S>(vol, a, c0 + 1.0f, c1 + 0.0f, c2 + 0.0f) -
                      sdf_sample<S>(vol, a, c0 - 1.0f, c1 - 0.0f, c2 - 0.0f)) / den;
    const float gy = (sdf_sample<S>(vol, a, c0 + 0.0f, c1 + 1.0f, c2 + 0.0f) -
                      sdf_sample<S>(vol, a, c0 - 0.0f, c1 - 1.0f, c2 - 0.0f)) / den;
    const float gz = (sdf_sample<S>(vol, a, c0 + 0.0f, c1 + 0.0f, c2 + 1.0f) -
                      sdf_sample<S>(vol, a, c0 - 0.0f, c1 - 0.0f, c2 - 1.0f)) / den;
    const float absD = D < 0.0f ? -D : D;
#pragma unroll
    for (int k = 0; k < 8; ++k) rec[k] = 0.0f;
    if (D < 1.0f && D > -1.0f) {
        const float px = pv[0], py = pv[1], pz = pv[2];
        rec[0] = absD < 0.02f ? 1.0f : 0.02f / absD;
        rec[1] = D;
        rec[2] = (gx * 0.0f + gy * (-pz)) + gz * py;
        rec[3] = (gx * pz + gy * 0.0f) + gz * (-px);
        rec[4] = (gx * (-py) + gy * px) + gz * 0.0f;
        rec[5] = (gx * 1.0f + gy * 0.0f) + gz * 0.0f;
        rec[6] = (gx * 0.0f + gy * 1.0f) + gz * 0.0f;
        rec[7] = (gx * 0.0f + gy * 0.0f) + gz * 1.0f;
    }
}

// slot of buffer index idx when it is y * W_l + W_l - 1 of some level l (the
// smallest such l), -1 otherwise
__device__ __forceinline__ int sdf_slot(const SdfArgs &a, int idx)
{
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const int w = a.w[l];
        if (w > 0 && idx < w * a.h[l] && idx % w == w - 1) return a.slot_base[l] + idx / w;
    }
    return -1;
}

// Eigen::LDLT<MatrixXd>(A).solve(b) (gFusion.cpp:918): symmetric pivoting on
// the largest remaining |diagonal|, D pseudo-inverse with threshold DBL_MIN,
// inner products summed in index order.  One thread; m (36, row-major), y, t
// (6 each) and tr (6) are LDS scratch, indexed at run time (pivots) without
// spilling to private memory.  On return y holds the solution.
__device__ void sdf_ldlt_solve(double *m, double *y, double *t, int *tr, const double (&b)[6])
{
#define M_(r, c) m[(r) * 6 + (c)]
    for (int k = 0; k < 6; ++k) {
        int big = k;
        double bv = fabs(M_(k, k));
        for (int i = k + 1; i < 6; ++i)
            if (fabs(M_(i, i)) > bv) {
                bv = fabs(M_(i, i));
                big = i;
            }
        tr[k] = big;
        if (big != k) {
            for (int j = 0; j < k; ++j) {
                const double u = M_(k, j);
                M_(k, j) = M_(big, j);
                M_(big, j) = u;
            }
            for (int i = big + 1; i < 6; ++i) {
                const double u = M_(i, k);
                M_(i, k) = M_(i, big);
                M_(i, big) = u;
            }
            {
                const double u = M_(k, k);
                M_(k, k) = M_(big, big);
                M_(big, big) = u;
            }
            for (int i = k + 1; i < big; ++i) {
                const double u = M_(i, k);
                M_(i, k) = M_(big, i);
                M_(big, i) = u;
            }
        }
        if (k > 0) {
            for (int j = 0; j < k; ++j) t[j] = M_(j, j) * M_(k, j);
            double sum = 0.0;
            for (int j = 0; j < k; ++j) sum += M_(k, j) * t[j];
            M_(k, k) -= sum;
            for (int i = k + 1; i < 6; ++i) {
                double u = 0.0;
                for (int j = 0; j < k; ++j) u += M_(i, j) * t[j];
                M_(i, k) -= u;
            }
        }
        const double akk = M_(k, k);
        const bool valid = fabs(akk) > 0.0;
        if (k == 0 && !valid) {
            // the whole diagonal is zero (or NaN): Eigen stops with the
            // identity permutation and the matrix as it is; the solve still runs
            for (int j = 0; j < 6; ++j) tr[j] = j;
            break;
        }
        if (valid)
            for (int i = k + 1; i < 6; ++i) M_(i, k) /= akk;
    }
    for (int i = 0; i < 6; ++i) y[i] = b[i];
    for (int k = 0; k < 6; ++k) {
        const double u = y[k];
        y[k] = y[tr[k]];
        y[tr[k]] = u;
    }
    for (int i = 0; i < 6; ++i) {
        double sum = y[i];
        for (int j = 0; j < i; ++j) sum -= M_(i, j) * y[j];
        y[i] = sum;
    }
    for (int i = 0; i < 6; ++i) y[i] = (fabs(M_(i, i)) > DBL_MIN) ? y[i] / M_(i, i) : 0.0;
    for (int i = 5; i >= 0; --i) {
        double sum = y[i];
        for (int j = i + 1; j < 6; ++j) sum -= M_(j, i) * y[j];
        y[i] = sum;
    }
    for (int k = 5; k >= 0; --k) {
        const double u = y[k];
        y[k] = y[tr[k]];
        y[tr[k]] = u;
    }
#undef M_
}

// Twist(xi).exp() (gFusion.cpp:810-833): closed-form SE(3) exponential of the
// twist held as float, evaluated in double, rounded to float; column-major.
__device__ void sdf_twist_exp(const double (&xi)[6], float out[16])
{
    const double w0 = (double)(float)xi[0], w1 = (double)(float)xi[1], w2 = (double)(float)xi[2];
    const double v[3] = {(double)(float)xi[3], (double)(float)xi[4], (double)(float)xi[5]};
    const double th2 = w0 * w0 + w1 * w1 + w2 * w2;
    double A, B, Cc;
    if (th2 < 1e-8) {
        A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
        Cc = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
    } else {
        const double th = sqrt(th2), s = sin(th), c = cos(th);
        A = s / th;
        B = (1.0 - c) / th2;
        Cc = (th - s) / (th2 * th);
    }
    const double W[3][3] = {{0.0, -w2, w1}, {w2, 0.0, -w0}, {-w1, w0, 0.0}};
    double W2[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) W2[r][c] = W[r][0] * W[0][c] + W[r][1] * W[1][c] + W[r][2] * W[2][c];
    double V[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const double I = (r == c) ? 1.0 : 0.0;
            out[c * 4 + r] = (float)(I + A * W[r][c] + B * W2[r][c]);
            V[r][c] = I + B * W[r][c] + Cc * W2[r][c];
        }
    for (int c = 0; c < 3; ++c) out[c * 4 + 3] = 0.0f;
    for (int r = 0; r < 3; ++r) out[12 + r] = (float)(V[r][0] * v[0] + V[r][1] * v[1] + V[r][2] * v[2]);
    out[15] = 1.0f;
}

// Grid barrier number `phase` (0-based) of this launch: returns false when
// the launch is being aborted.
__device__ __forceinline__ bool sdf_grid_barrier(DevState *st, unsigned phase, unsigned nblk, unsigned spin_limit,
                                                 int *sh_abort)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // release this block's partials and slot records
        atomicAdd(&st->sdf_bar, 1u);
        const unsigned target = (phase + 1u) * nblk;
        unsigned n = 0;
        int ab = 0;
        for (;;) {
            if (__hip_atomic_load(&st->sdf_bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
            if (__hip_atomic_load(&st->sdf_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ab = 1;
                break;
            }
            if (++n > spin_limit) {
                __hip_atomic_store(&st->sdf_abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ab = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        *sh_abort = ab;
    }
    __syncthreads();
    __threadfence();  // acquire: the other blocks' partials and records
    return *sh_abort == 0;
}

template <class S>
__global__ __launch_bounds__(kSdfThreads) void k_sdf_track(SdfArgs a, DevState *__restrict__ st)
{
    __shared__ float wsum[kSdfThreads / 64][kSdfTerms];
    __shared__ double wgsum[kSdfThreads / 64][8][kSdfTerms];
    __shared__ double gsum[8][kSdfTerms];
    __shared__ float T[16], prev[16];
    __shared__ double result[6], result_prev[6];
    __shared__ double solve_m[36], solve_y[6], solve_t[6];
    __shared__ int solve_tr[6];
    __shared__ int converged, sh_abort;

    const int blk = blockIdx.x, nblk = gridDim.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const typename S::V *__restrict__ vol = (const typename S::V *)a.tsdf;

    // the frame loop's dilation of the brick flags (independent of the
    // tracking below, which reads the TSDF only): four coarse cells per
    // wavefront per round
    if (a.dilate) {
        const int waves = nblk * (kSdfThreads / 64);
        const int nc = a.bricks.NC[0] * a.bricks.NC[1] * a.bricks.NC[2];
        for (int i0 = blk * (kSdfThreads / 64) + wave; i0 < nc; i0 += 4 * waves) brick_dilate_cells4(a.bricks, i0, waves);
    }

    // which entries to run (uniform over the grid: the state read here was
    // written by earlier launches)
    int entries = 1;
    if (a.mode != kSdfTrack) {
        const int n = st->lib_n;
        if (a.mode == kSdfRecoverIfLost && st->tracked != 0) return;
        if (n <= 0) {
            if (a.mode == kSdfRecover && blk == 0 && tid == 0) st->sdf_tracked = 0;
            return;
        }
        entries = n;
    }
    if (tid == 0) sh_abort = 0;
    if (tid < 16) prev[tid] = st->pose_eig[tid];
    __syncthreads();

    unsigned phase = 0;
    int found = 0, iters = 0;
    bool aborted = false;
    for (int e = 0; e < entries && !found && !aborted; ++e) {
        if (tid < 6) {
            result[tid] = 0.0;
            result_prev[tid] = 0.0;
        }
        int tracked = 0;
        for (int level = 2; level >= 0 && !aborted; --level) {
            const int w = a.w[level], h = a.h[level];
            const float4 *__restrict__ vert = a.vert[level];
            for (int it = 0; it < a.iters[level]; ++it) {
                __syncthreads();
                if (tid == 0) {
                    // camToWorld = exp(Twist(result)) * m_pose_eig_prev (gFusion.cpp:866)
                    double r[6];
                    for (int k = 0; k < 6; ++k) r[k] = result[k];
                    float E[16], P[16];
                    sdf_twist_exp(r, E);
                    for (int k = 0; k < 16; ++k) P[k] = prev[k];
                    mul4(E, P, E);
                    for (int k = 0; k < 16; ++k) T[k] = E[k];
                }
                __syncthreads();
                float Tm[16];
#pragma unroll
                for (int k = 0; k < 16; ++k) Tm[k] = T[k];

                // this block's rows: y = g + 8 (j + G8 m), g = blk % 8
                float acc[kSdfTerms];
#pragma unroll
                for (int k = 0; k < kSdfTerms; ++k) acc[k] = 0.0f;
                const int g = blk & 7, G8 = nblk >> 3;
                for (int y = g + 8 * (blk >> 3); y < h; y += 8 * G8) {
                    for (int x = tid; x < w; x += kSdfThreads) {
                        const int idx = y * w + x;
                        float rec[8];
                        if (x < w - 1) {
                            sdf_record<S>(vol, a, Tm, vert[idx], rec);
                            const int sl = sdf_slot(a, idx);
                            if (sl >= 0) {
                                float4 *d = reinterpret_cast<float4 *>(a.slots + 8 * sl);
                                d[0] = make_float4(rec[0], rec[1], rec[2], rec[3]);
                                d[1] = make_float4(rec[4], rec[5], rec[6], rec[7]);
                            }
                        } else {
                            // the record an earlier dispatch left at this index
                            const int sl = sdf_slot(a, idx);
                            const float4 *d = reinterpret_cast<const float4 *>(a.slots + 8 * sl);
                            const float4 r0 = d[0], r1 = d[1];
                            rec[0] = r0.x; rec[1] = r0.y; rec[2] = r0.z; rec[3] = r0.w;
                            rec[4] = r1.x; rec[5] = r1.y; rec[6] = r1.z; rec[7] = r1.w;
                        }
                        // reduceSDF.cs:50-101
                        const float hh = rec[0], D = rec[1];
                        acc[0] += D * D;
#pragma unroll
                        for (int k = 0; k < 6; ++k) acc[1 + k] += hh * rec[2 + k] * D;
                        acc[0] += D * D;
                        int q = 7;
#pragma unroll
                        for (int p = 0; p < 6; ++p)
#pragma unroll
                            for (int s2 = p; s2 < 6; ++s2) acc[q++] += hh * rec[2 + p] * rec[2 + s2];
                    }
                }
                wave_sums_store(acc, wsum[wave]);
                __syncthreads();
                float *part = a.partials + (size_t)(phase & 1u) * nblk * 32;
                if (tid < kSdfTerms)
                    part[blk * 32 + tid] = (wsum[0][tid] + wsum[1][tid]) + (wsum[2][tid] + wsum[3][tid]);
                if (!sdf_grid_barrier(st, phase, (unsigned)nblk, a.spin_limit, &sh_abort)) {
                    aborted = true;
                    break;
                }
                ++phase;

                // FP64 sums per row group, the same fixed order in every block
                {
                    double ps[kSdfTerms];
#pragma unroll
                    for (int k = 0; k < kSdfTerms; ++k) ps[k] = 0.0;
                    for (int pb = tid; pb < nblk; pb += kSdfThreads) {
                        const float4 *rp = reinterpret_cast<const float4 *>(part + pb * 32);
                        float4 qv[kSdfTerms / 4];
#pragma unroll
                        for (int m = 0; m < kSdfTerms / 4; ++m) qv[m] = rp[m];
#pragma unroll
                        for (int m = 0; m < kSdfTerms / 4; ++m) {
                            ps[4 * m + 0] += (double)qv[m].x;
                            ps[4 * m + 1] += (double)qv[m].y;
                            ps[4 * m + 2] += (double)qv[m].z;
                            ps[4 * m + 3] += (double)qv[m].w;
                        }
                    }
#pragma unroll
                    for (int k = 0; k < kSdfTerms; ++k) {
                        ps[k] += __shfl_xor(ps[k], 8, 64);
                        ps[k] += __shfl_xor(ps[k], 16, 64);
                        ps[k] += __shfl_xor(ps[k], 32, 64);
                    }
                    if (lane < 8) {
#pragma unroll
                        for (int k = 0; k < kSdfTerms; ++k) wgsum[wave][lane][k] = ps[k];
                    }
                }
                __syncthreads();
                if (tid < 8 * kSdfTerms) {
                    const int gg = tid / kSdfTerms, k = tid % kSdfTerms;
                    gsum[gg][k] = (wgsum[0][gg][k] + wgsum[1][gg][k]) + (wgsum[2][gg][k] + wgsum[3][gg][k]);
                }
                __syncthreads();
                if (blk == 0) {
                    // the 8x32 partials in the reduceSDF.cs layout; [28] counts
                    // every pixel of the group's rows
                    const int gg = tid >> 5, k = tid & 31;
                    const int rows = (h > gg) ? (h - gg + 7) / 8 : 0;
                    float val = 0.0f;
                    if (k < kSdfTerms) val = (float)gsum[gg][k];
                    else if (k == kSdfTerms) val = (float)((double)rows * (double)w);
                    st->sdf_partials[tid] = val;
                }
                if (tid == 0) {
                    // getSDFReduction (gFusion.cpp:1195-1232) + the step of :878-931
                    double v[kSdfTerms], count = 0.0;
                    for (int k = 0; k < kSdfTerms; ++k) v[k] = 0.0;
                    if (a.rule == 1) {
                        for (int gg = 0; gg < 8; ++gg) {
                            for (int k = 0; k < kSdfTerms; ++k) v[k] += gsum[gg][k];
                            count += (double)((h > gg) ? (h - gg + 7) / 8 : 0) * (double)w;
                        }
                    } else {
                        for (int k = 0; k < kSdfTerms; ++k) v[k] = gsum[7][k];
                        count = (double)((h > 7) ? (h - 7 + 7) / 8 : 0) * (double)w;
                    }
                    const float ae = sqrtf((float)v[0] / (float)count);
                    double bb[6];
                    {
                        int q = 7;
#pragma unroll
                        for (int p = 0; p < 6; ++p)
#pragma unroll
                            for (int s2 = p; s2 < 6; ++s2) {
                                solve_m[p * 6 + s2] = solve_m[s2 * 6 + p] = (double)(float)v[q];
                                ++q;
                            }
                    }
                    double mx = solve_m[0];
                    for (int k = 1; k < 36; ++k)
                        if (solve_m[k] > mx) mx = solve_m[k];
                    const double scaling = 1.0 / mx;
                    for (int k = 0; k < 36; ++k) solve_m[k] *= scaling;
#pragma unroll
                    for (int k = 0; k < 6; ++k) bb[k] = (double)(float)v[1 + k] * scaling;
                    for (int k = 0; k < 6; ++k) solve_m[k * 6 + k] += (double)it * 0.01;
                    sdf_ldlt_solve(solve_m, solve_y, solve_t, solve_tr, bb);
                    double cn = 0.0, r[6];
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        r[k] = result[k] - solve_y[k];
                        const double d = r[k] - result_prev[k];
                        cn += d * d;
                    }
                    cn = sqrt(cn);
                    for (int k = 0; k < 6; ++k) {
                        result[k] = r[k];
                        result_prev[k] = r[k];
                    }
                    const bool conv = ((double)ae != 0.0 && (double)ae < 1e-3) || cn < 1e-3;
                    converged = conv ? 1 : 0;
                    if (blk == 0) {
                        st->sdf_ae = ae;
                        if (iters < 64) {
                            double *L = st->sdf_log[iters];
                            L[0] = level;
                            L[1] = it;
                            L[2] = (double)ae;
                            L[3] = cn;
                            for (int k = 0; k < 6; ++k) L[4 + k] = r[k];
                        }
                    }
                }
                __syncthreads();
                ++iters;
                if (converged) {
                    tracked = 1;
                    break;
                }
            }
        }
        if (aborted) break;
        // end of TrackSDF (gFusion.cpp:988-1012), written by block 0
        if (blk == 0 && tid == 0) {
            double r[6], sum = 0.0;
            for (int k = 0; k < 6; ++k) {
                r[k] = result[k];
                sum += r[k];
            }
            if (sum != sum)
                for (int k = 0; k < 6; ++k) r[k] = 0.0;
            if (tracked) {
                float E[16], P[16];
                sdf_twist_exp(r, E);
                for (int k = 0; k < 16; ++k) P[k] = prev[k];
                mul4(E, P, E);
                for (int k = 0; k < 16; ++k) {
                    st->pose_eig[k] = E[k];
                    st->pose[k] = E[k];
                }
                st->pose[3] = st->pose[7] = st->pose[11] = 0.0f;
                st->pose[15] = 1.0f;
                for (int k = 0; k < 6; ++k) st->cum_twist[k] += r[k];
            } else if (a.mode == kSdfTrack) {
                // m_pose = oldPose: unchanged (TrackSDF itself does not move m_pose
                // before the end); m_pose_eig = m_pose_eig_prev: unchanged
            } else {
                // recoverPose set m_pose to the library entry before TrackSDF,
                // and a failed TrackSDF restores that
                for (int k = 0; k < 16; ++k) st->pose[k] = a.library[16 * e + k];
            }
        }
        found = tracked;
    }

    if (blk == 0 && tid == 0) {
        st->sdf_iters = iters;
        if (aborted) {
            st->sdf_error = 1;
        } else {
            st->sdf_tracked = found;
            if (a.mode == kSdfTrack) {
                st->tracked = found;
                if (a.frame) {
                    const int go = found || st->reset;
                    st->integrate_now = go;
                    if (go && st->counter > 2) st->reset = 0;
                    st->counter += 1;
                }
            }
        }
    }
    // exit ticket: the last block out clears the barrier state for the next launch
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        const unsigned t = atomicAdd(&st->sdf_exit, 1u);
        if (t == (unsigned)(nblk - 1)) {
            st->sdf_bar = 0u;
            st->sdf_abort = 0;
            __threadfence();
            st->sdf_exit = 0u;
        }
    }
}

hipError_t launch_sdf_track(const SdfArgs &a, int storage, DevState *st, int nblk, hipStream_t s)
{
    if (nblk < 8 || (nblk & 7)) return hipErrorInvalidValue;
    switch (storage) {
    case KF_STORAGE_RG16I: hipLaunchKernelGGL(k_sdf_track<StoreRG16I>, dim3(nblk), dim3(kSdfThreads), 0, s, a, st); break;
    case KF_STORAGE_F32: hipLaunchKernelGGL(k_sdf_track<StoreF32>, dim3(nblk), dim3(kSdfThreads), 0, s, a, st); break;
    case KF_STORAGE_F16: hipLaunchKernelGGL(k_sdf_track<StoreF16>, dim3(nblk), dim3(kSdfThreads), 0, s, a, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace kf
