/*
 * kf_oracle.c -- CPU oracle for the per-frame KinectFusion loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the algorithm the
 * reference runs as GLSL compute shaders plus host C++ (philipNoonan/
 * oglFusionKinectV2).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline; the
 * product library never links or calls it.
 *
 * It is compiled twice (see Makefile):
 *   libkforacle_f32.so  R = float : reproduces the shader/host FP32 semantics
 *                                   (no FMA contraction, IEEE div/sqrt)
 *   libkforacle_f64.so  R = double: the "FP64 oracle" for the precision variant
 *
 * Every function cites the reference file:line it follows.  Matrices are 16
 * values, column-major: M[c*4 + r] is glm's M[c][r].
 *
 * Parity pinning (tests/test_oracle_*.py): closed forms, special cases, brute
 * force on small inputs, numpy.linalg for the 6x6 solve, numpy.float16 for the
 * half conversion; and the reference's own compute shaders, run by Mesa's
 * llvmpipe on identical inputs (tests/test_oracle_reference_shaders.py):
 * depthToVert, reduce and first-frame integrate bit-exact, track's result
 * codes and Jacobians bit-exact, vertToNorm / raycast within the ulps a GL
 * compiler may spend on FMA contraction, the driver's mip filter the box mean
 * up to its bilinear-weight rounding, and the frame loop within 1 mm over 8
 * frames.  The host-side arithmetic of glm (eulerAngleXYZ, inverse) and Eigen
 * (JacobiSVD, ldlt, the matrix exponential) cannot be built here and stays
 * "parity unpinned" at the ulp level: it is pinned to its mathematical
 * contract only.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#if defined(KFO_F64)
typedef double R;
#define KC(x) x
#define KSQRT sqrt
#define KFLOOR floor
#define KMODF modf
#else
typedef float R;
#define KC(x) x##f
#define KSQRT sqrtf
#define KFLOOR floorf
#define KMODF modff
#endif

#define KFO_EXPORT __attribute__((visibility("default")))

/* TSDF storage modes */
enum { KFO_RG16I = 0, KFO_F32X2 = 1, KFO_F16X2 = 2, KFO_F64X2 = 3 };
/* reduction-combine rules for the 8x32 partials */
enum { KFO_RULE_LAST_ROW = 0, KFO_RULE_SUM_ALL = 1 };

KFO_EXPORT int kfo_real_size(void) { return (int)sizeof(R); }

/* ------------------------------------------------------------------------- */
/* IEEE binary16 conversion (round to nearest even), own implementation.      */
/* ------------------------------------------------------------------------- */
KFO_EXPORT uint16_t kfo_f2h(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) /* inf / nan */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) /* >= 65520 rounds to inf */
        return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u) { /* below 2^-14: subnormal half or zero */
        if (ax <= 0x33000000u) /* <= 2^-25 rounds to zero (tie to even) */
            return (uint16_t)sign;
        uint32_t e = ax >> 23;
        uint32_t m = (ax & 0x7fffffu) | 0x800000u;
        uint32_t shift = 126u - e; /* 14..24 */
        uint32_t q = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1u);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t e = (ax >> 23) - 127u + 15u;
    uint32_t m = ax & 0x7fffffu;
    uint32_t q = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    return (uint16_t)(sign | q);
}

KFO_EXPORT float kfo_h2f(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t m = (uint32_t)h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {
            float v = (float)m * 5.9604644775390625e-8f; /* m * 2^-24, exact */
            memcpy(&x, &v, 4);
            x |= sign;
        }
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e - 15u + 127u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* ------------------------------------------------------------------------- */
/* GLSL helpers of the shaders: dot / rotate / opMul / mat*vec / min / max      */
/* ------------------------------------------------------------------------- */
static inline R dot3(R ax, R ay, R az, R bx, R by, R bz)
{
    return (ax * bx + ay * by) + az * bz;
}

/* rotate(M, V): shaders/integrate.cs:21-27 (same in track.cs:37-43, raycast.cs:43-49) */
static inline void rotate3(const R *M, R vx, R vy, R vz, R out[3])
{
    for (int i = 0; i < 3; ++i)
        out[i] = dot3(M[0 * 4 + i], M[1 * 4 + i], M[2 * 4 + i], vx, vy, vz);
}

/* opMul(M, v): shaders/integrate.cs:29-35 */
static inline void opmul3(const R *M, R vx, R vy, R vz, R out[3])
{
    for (int i = 0; i < 3; ++i)
        out[i] = dot3(M[0 * 4 + i], M[1 * 4 + i], M[2 * 4 + i], vx, vy, vz) + M[3 * 4 + i];
}

/* GLSL mat4 * vec4 (column combination, summed left to right) */
static inline void matvec4(const R *M, const R v[4], R out[4])
{
    for (int i = 0; i < 4; ++i)
        out[i] = ((M[0 * 4 + i] * v[0] + M[1 * 4 + i] * v[1]) + M[2 * 4 + i] * v[2]) + M[3 * 4 + i] * v[3];
}

/* GLSL min/max: min(x,y) = y < x ? y : x ; max(x,y) = x < y ? y : x */
static inline R gmin(R x, R y) { return y < x ? y : x; }
static inline R gmax(R x, R y) { return x < y ? y : x; }
static inline int imin(int x, int y) { return y < x ? y : x; }
static inline int imax(int x, int y) { return x < y ? y : x; }

/* float -> int conversion of a GLSL ivec() cast; NaN / out-of-range map to
 * INT_MIN so that a following image load is out of bounds and reads zero. */
static inline int f2i(R v)
{
    if (!(v > KC(-2147483648.0) && v < KC(2147483648.0))) return INT32_MIN;
    return (int)v;
}

/* ------------------------------------------------------------------------- */
/* Matrices (glm semantics): include/gFusion.h:451-470, glm mat4 product      */
/* ------------------------------------------------------------------------- */

/* getCameraMatrix: include/gFusion.h:461-470 */
KFO_EXPORT void kfo_camera_matrix(const R cam[4], R K[16])
{
    memset(K, 0, 16 * sizeof(R));
    K[0] = cam[0];
    K[5] = cam[1];
    K[2 * 4 + 0] = cam[2];
    K[2 * 4 + 1] = cam[3];
    K[10] = KC(1.0);
    K[15] = KC(1.0);
}

/* getInverseCameraMatrix: include/gFusion.h:451-459 */
KFO_EXPORT void kfo_inverse_camera_matrix(const R cam[4], R iK[16])
{
    memset(iK, 0, 16 * sizeof(R));
    iK[0] = KC(1.0) / cam[0];
    iK[5] = KC(1.0) / cam[1];
    iK[2 * 4 + 0] = -cam[2] / cam[0];
    iK[2 * 4 + 1] = -cam[3] / cam[1];
    iK[10] = KC(1.0);
    iK[15] = KC(1.0);
}

/* glm mat4 * mat4: result column j = A[0]*B[j][0] + A[1]*B[j][1] + A[2]*B[j][2] + A[3]*B[j][3] */
KFO_EXPORT void kfo_mul4(const R *A, const R *B, R *out)
{
    R t[16];
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            t[j * 4 + i] = ((A[0 * 4 + i] * B[j * 4 + 0] + A[1 * 4 + i] * B[j * 4 + 1]) +
                            A[2 * 4 + i] * B[j * 4 + 2]) + A[3 * 4 + i] * B[j * 4 + 3];
    memcpy(out, t, sizeof(t));
}

/* General 4x4 inverse (stands in for glm::inverse at gFusion.cpp:747, 1242).
 * 2x2 sub-determinant (Laplace) expansion.  a(r,c) = M[c*4+r]. */
KFO_EXPORT void kfo_inverse4(const R *M, R *out)
{
#define A(r, c) M[(c) * 4 + (r)]
    R s0 = A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1);
    R s1 = A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2);
    R s2 = A(0, 0) * A(1, 3) - A(1, 0) * A(0, 3);
    R s3 = A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2);
    R s4 = A(0, 1) * A(1, 3) - A(1, 1) * A(0, 3);
    R s5 = A(0, 2) * A(1, 3) - A(1, 2) * A(0, 3);
    R c5 = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3);
    R c4 = A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3);
    R c3 = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2);
    R c2 = A(2, 0) * A(3, 3) - A(3, 0) * A(2, 3);
    R c1 = A(2, 0) * A(3, 2) - A(3, 0) * A(2, 2);
    R c0 = A(2, 0) * A(3, 1) - A(3, 0) * A(2, 1);
    R det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    R id = KC(1.0) / det;
    R B[16];
#define Bs(r, c, v) B[(c) * 4 + (r)] = (v) * id
    Bs(0, 0, A(1, 1) * c5 - A(1, 2) * c4 + A(1, 3) * c3);
    Bs(0, 1, -A(0, 1) * c5 + A(0, 2) * c4 - A(0, 3) * c3);
    Bs(0, 2, A(3, 1) * s5 - A(3, 2) * s4 + A(3, 3) * s3);
    Bs(0, 3, -A(2, 1) * s5 + A(2, 2) * s4 - A(2, 3) * s3);
    Bs(1, 0, -A(1, 0) * c5 + A(1, 2) * c2 - A(1, 3) * c1);
    Bs(1, 1, A(0, 0) * c5 - A(0, 2) * c2 + A(0, 3) * c1);
    Bs(1, 2, -A(3, 0) * s5 + A(3, 2) * s2 - A(3, 3) * s1);
    Bs(1, 3, A(2, 0) * s5 - A(2, 2) * s2 + A(2, 3) * s1);
    Bs(2, 0, A(1, 0) * c4 - A(1, 1) * c2 + A(1, 3) * c0);
    Bs(2, 1, -A(0, 0) * c4 + A(0, 1) * c2 - A(0, 3) * c0);
    Bs(2, 2, A(3, 0) * s4 - A(3, 1) * s2 + A(3, 3) * s0);
    Bs(2, 3, -A(2, 0) * s4 + A(2, 1) * s2 - A(2, 3) * s0);
    Bs(3, 0, -A(1, 0) * c3 + A(1, 1) * c1 - A(1, 2) * c0);
    Bs(3, 1, A(0, 0) * c3 - A(0, 1) * c1 + A(0, 2) * c0);
    Bs(3, 2, -A(3, 0) * s3 + A(3, 1) * s1 - A(3, 2) * s0);
    Bs(3, 3, A(2, 0) * s3 - A(2, 1) * s1 + A(2, 2) * s0);
#undef Bs
#undef A
    memcpy(out, B, sizeof(B));
}

/* ------------------------------------------------------------------------- */
/* Pyramid: gFusion.cpp:496-547, depthToVert.cs:33-72, vertToNorm.cs:9-57     */
/* ------------------------------------------------------------------------- */

/* glGenerateMipmap of the R32F depth texture (gFusion.cpp:508): 2x2 box mean,
 * zeros included.  The filter is the driver's choice; llvmpipe's agrees with
 * this mean to within 2.5e-4 of the block's range (its bilinear weights are
 * rounded), which the tests pin. */
KFO_EXPORT void kfo_mip_depth(const R *src, int w, int h, R *dst)
{
    int w2 = w >> 1, h2 = h >> 1;
    for (int y = 0; y < h2; ++y)
        for (int x = 0; x < w2; ++x) {
            const R *p = src + (2 * y) * w + 2 * x;
            dst[y * w2 + x] = ((p[0] + p[1]) + (p[w] + p[w + 1])) * KC(0.25);
        }
}

/* depthToVert.cs:56 with invK = getInverseCameraMatrix(camPams / 2^l) (gFusion.cpp:517) */
KFO_EXPORT void kfo_depth_to_vertex(const R *depth, int w, int h, const R cam_l[4], R *vert)
{
    R iK[16];
    kfo_inverse_camera_matrix(cam_l, iK);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            R r[3];
            rotate3(iK, (R)x, (R)y, KC(1.0), r);
            R s = depth[y * w + x] / KC(1000.0);
            R *o = vert + 4 * (y * w + x);
            o[0] = s * r[0];
            o[1] = s * r[1];
            o[2] = s * r[2];
            o[3] = KC(0.0);
        }
}

/* vertToNorm.cs:9-57.  Left/up use max(pix-1, 0) on an unsigned pixel index,
 * which wraps to -1 at the border; down reads (x-1, y+1).  Out-of-bounds image
 * loads return zero, so border pixels become the invalid sentinel (2,0,0,0). */
KFO_EXPORT void kfo_vertex_to_normal(const R *vert, int w, int h, R *norm)
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int lx = x - 1, ly = y;                         /* left  */
            int rx = imin(x + 1, w - 1), ry = y;            /* right */
            int ux = x, uy = y - 1;                         /* up    */
            int dx = x - 1, dy = imin(y + 1, h - 1);        /* down  */
            R L[3] = {0, 0, 0}, Rt[3] = {0, 0, 0}, U[3] = {0, 0, 0}, D[3] = {0, 0, 0};
            if (lx >= 0) memcpy(L, vert + 4 * (ly * w + lx), 3 * sizeof(R));
            memcpy(Rt, vert + 4 * (ry * w + rx), 3 * sizeof(R));
            if (uy >= 0) memcpy(U, vert + 4 * (uy * w + ux), 3 * sizeof(R));
            if (dx >= 0) memcpy(D, vert + 4 * (dy * w + dx), 3 * sizeof(R));
            R *o = norm + 4 * (y * w + x);
            if (L[2] == 0 || Rt[2] == 0 || U[2] == 0 || D[2] == 0) {
                o[0] = KC(2.0); o[1] = 0; o[2] = 0; o[3] = 0;
                continue;
            }
            R dxv[3] = {Rt[0] - L[0], Rt[1] - L[1], Rt[2] - L[2]};
            R dyv[3] = {D[0] - U[0], D[1] - U[1], D[2] - U[2]};
            /* cross(dyv, dxv) */
            R c0 = dyv[1] * dxv[2] - dyv[2] * dxv[1];
            R c1 = dyv[2] * dxv[0] - dyv[0] * dxv[2];
            R c2 = dyv[0] * dxv[1] - dyv[1] * dxv[0];
            R len = KSQRT(dot3(c0, c1, c2, c0, c1, c2));
            o[0] = c0 / len; o[1] = c1 / len; o[2] = c2 / len; o[3] = KC(1.0);
        }
}

/* ------------------------------------------------------------------------- */
/* ICP: gFusion.cpp:742-777 + track.cs:47-163                                  */
/* ------------------------------------------------------------------------- */

/* view = getCameraMatrix(camPams) * inverse(m_pose)   (gFusion.cpp:747) */
KFO_EXPORT void kfo_track_view(const R pose[16], const R cam[4], R view[16])
{
    R K[16], ip[16];
    kfo_camera_matrix(cam, K);
    kfo_inverse4(pose, ip);
    kfo_mul4(K, ip, view);
}

static void track_pixel(const R *vert, const R *norm, int w, int x, int y,
                        const R *refV, const R *refN, int rw, int rh,
                        const R *T, const R *view, R dist_th, R normal_th,
                        int *res, R *err, R *J)
{
    const R *n = norm + 4 * (y * w + x);
    if (n[0] == KC(2.0)) { *res = -1; return; }
    const R *v = vert + 4 * (y * w + x);
    R v4[4] = {v[0], v[1], v[2], KC(1.0)};
    R pv[4], pp[4];
    matvec4(T, v4, pv);     /* projectedVertex = Ttrack * vec4(v, 1) */
    matvec4(view, pv, pp);  /* projectedPos = view * projectedVertex */
    R px = pp[0] / pp[2] + KC(0.5);
    R py = pp[1] / pp[2] + KC(0.5);
    if (px < 0 || px > (R)(rw - 1) || py < 0 || py > (R)(rh - 1)) { *res = -2; return; }
    int ix = f2i(px), iy = f2i(py);
    R rn[3] = {0, 0, 0}, rv[3] = {0, 0, 0};
    if (ix >= 0 && ix < rw && iy >= 0 && iy < rh) {
        memcpy(rn, refN + 4 * (iy * rw + ix), 3 * sizeof(R));
        memcpy(rv, refV + 4 * (iy * rw + ix), 3 * sizeof(R));
    }
    if (rn[0] == KC(-2.0)) { *res = -3; return; }
    R diff[3] = {rv[0] - pv[0], rv[1] - pv[1], rv[2] - pv[2]};
    R pn[3];
    rotate3(T, n[0], n[1], n[2], pn);
    if (KSQRT(dot3(diff[0], diff[1], diff[2], diff[0], diff[1], diff[2])) > dist_th) { *res = -4; return; }
    if (dot3(pn[0], pn[1], pn[2], rn[0], rn[1], rn[2]) < normal_th) { *res = -5; return; }
    *res = 1;
    *err = dot3(rn[0], rn[1], rn[2], diff[0], diff[1], diff[2]);
    J[0] = rn[0]; J[1] = rn[1]; J[2] = rn[2];
    /* cross(projectedVertex.xyz, referenceNormal) */
    J[3] = pv[1] * rn[2] - pv[2] * rn[1];
    J[4] = pv[2] * rn[0] - pv[0] * rn[2];
    J[5] = pv[0] * rn[1] - pv[1] * rn[0];
}

/* track.cs main over a (w x h) level; results per pixel: result, error, J[6].
 * error/J are written only for result == 1 (other slots keep their values). */
KFO_EXPORT void kfo_track(const R *vert, const R *norm, int w, int h,
                          const R *refV, const R *refN, int rw, int rh,
                          const R pose[16], const R view[16], R dist_th, R normal_th,
                          int *result, R *error, R *J)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int i = y * w + x;
            track_pixel(vert, norm, w, x, y, refV, refN, rw, rh, pose, view, dist_th, normal_th,
                        result + i, error + i, J + 6 * i);
        }
}

/* reduce.cs:27-127 with the dispatch of gFusion.cpp:779-797: 8 work groups of
 * 112 threads; WG g takes rows y = g (mod 8), thread s takes columns x = s
 * (mod 112), each thread sums serially, then threads 0..31 add the 112 rows
 * serially.  out[g*32 + k]. */
KFO_EXPORT void kfo_reduce(const int *result, const R *error, const R *J, int w, int h, R out[256])
{
    R (*S)[112][32] = (R (*)[112][32])malloc(sizeof(R) * 8 * 112 * 32);
#pragma omp parallel for schedule(static)
    for (int gs = 0; gs < 8 * 112; ++gs) {
        int g = gs / 112, s = gs % 112;
        R sums[32];
        for (int k = 0; k < 32; ++k) sums[k] = 0;
        for (int y = g; y < h; y += 8)
            for (int x = s; x < w; x += 112) {
                int i = y * w + x;
                int r = result[i];
                if (r < 1) {
                    if (r == -4) sums[28] += 1;
                    if (r == -5) sums[28] += 1;
                    if (r > -4) sums[28] += 1;
                    continue;
                }
                R e = error[i];
                const R *j = J + 6 * i;
                sums[0] += e * e;
                for (int k = 0; k < 6; ++k) sums[k + 1] += e * j[k];
                int q = 7;
                for (int a = 0; a < 6; ++a)
                    for (int b = a; b < 6; ++b) sums[q++] += j[a] * j[b];
                sums[28] += KC(1.0);
            }
        memcpy(S[g][s], sums, sizeof(sums));
    }
    for (int g = 0; g < 8; ++g)
        for (int k = 0; k < 32; ++k) {
            R acc = S[g][0][k];
            for (int s = 1; s < 112; ++s) acc += S[g][s][k];
            out[g * 32 + k] = acc;
        }
    free(S);
}

/* makeJTJ: include/gFusion.h:472-487 -- 21 upper-triangle values -> symmetric 6x6 */
static void make_jtj(const R *v, R C[36])
{
    int q = 0;
    for (int a = 0; a < 6; ++a)
        for (int b = a; b < 6; ++b) {
            C[a * 6 + b] = v[q];
            C[b * 6 + a] = v[q];
            ++q;
        }
}

/* getReduction: gFusion.cpp:1127-1168.  Rule LAST_ROW reproduces the loop at
 * :1143-1149 that assigns (not adds) rows 1..7 into row 0, so only WG 7's
 * partial survives.  Rule SUM_ALL adds rows 0..7 in order. */
KFO_EXPORT void kfo_get_reduction(const R partials[256], int rule, R b[6], R C[36], R *ae)
{
    R row[32];
    if (rule == KFO_RULE_SUM_ALL) {
        for (int k = 0; k < 32; ++k) {
            R acc = partials[k];
            for (int g = 1; g < 8; ++g) acc += partials[g * 32 + k];
            row[k] = acc;
        }
    } else {
        memcpy(row, partials + 7 * 32, 32 * sizeof(R));
    }
    for (int k = 0; k < 6; ++k) b[k] = row[1 + k];
    make_jtj(row + 7, C);
    *ae = KSQRT(row[0] / row[28]);
}

/* Eigen::JacobiSVD<MatrixXd>(C).solve(b) for the symmetric 6x6 of
 * gFusion.cpp:696-697: minimum-norm least-squares solution, singular values
 * below 6*eps*sigma_max treated as zero (Eigen's default threshold).  For a
 * symmetric matrix the SVD follows from the eigen-decomposition (cyclic
 * Jacobi rotations here). */
KFO_EXPORT void kfo_solve6(const double Cin[36], const double b[6], double x[6])
{
    double A[6][6], V[6][6];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            A[i][j] = 0.5 * (Cin[i * 6 + j] + Cin[j * 6 + i]);
            V[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < 6; ++i) {
            diag += A[i][i] * A[i][i];
            for (int j = i + 1; j < 6; ++j) off += A[i][j] * A[i][j];
        }
        if (off == 0.0 || off <= 1e-34 * diag) break;
        for (int p = 0; p < 5; ++p)
            for (int q = p + 1; q < 6; ++q) {
                double apq = A[p][q];
                if (apq == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 6; ++k) { /* A <- A J */
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 6; ++k) { /* A <- J^T A */
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 6; ++k) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double smax = 0.0;
    for (int i = 0; i < 6; ++i) smax = fmax(smax, fabs(A[i][i]));
    double thr = fmax(smax * 6.0 * DBL_EPSILON, DBL_MIN);
    for (int k = 0; k < 6; ++k) x[k] = 0.0;
    for (int i = 0; i < 6; ++i) {
        double lam = A[i][i];
        if (fabs(lam) < thr) continue;
        double proj = 0.0;
        for (int k = 0; k < 6; ++k) proj += V[k][i] * b[k];
        proj /= lam;
        for (int k = 0; k < 6; ++k) x[k] += V[k][i] * proj;
    }
}

/* glm::eulerAngleXYZ(x3, x4, x5) evaluated in double = Rx(x3)*Ry(x4)*Rz(x5),
 * translation (x0, x1, x2), cast to R; then m_pose = delta * m_pose
 * (gFusion.cpp:700-705).  No re-orthonormalisation. */
KFO_EXPORT void kfo_pose_update(const double x[6], R pose[16])
{
    double ca = cos(x[3]), sa = sin(x[3]);
    double cb = cos(x[4]), sb = sin(x[4]);
    double cc = cos(x[5]), sc = sin(x[5]);
    double Rm[3][3] = {
        {cb * cc, -cb * sc, sb},
        {sa * sb * cc + ca * sc, -sa * sb * sc + ca * cc, -sa * cb},
        {-ca * sb * cc + sa * sc, ca * sb * sc + sa * cc, ca * cb},
    };
    R d[16];
    for (int c = 0; c < 3; ++c) {
        for (int r = 0; r < 3; ++r) d[c * 4 + r] = (R)Rm[r][c];
        d[c * 4 + 3] = 0;
    }
    d[12] = (R)x[0]; d[13] = (R)x[1]; d[14] = (R)x[2]; d[15] = KC(1.0);
    kfo_mul4(d, pose, pose);
}

/* ------------------------------------------------------------------------- */
/* TSDF storage                                                               */
/* ------------------------------------------------------------------------- */
KFO_EXPORT int kfo_voxel_bytes(int mode)
{
    switch (mode) {
    case KFO_RG16I: return 4;
    case KFO_F32X2: return 8;
    case KFO_F16X2: return 4;
    case KFO_F64X2: return 16;
    }
    return 0;
}

/* vs(): integrate.cs:37-43 -- (value, weight) as R */
static inline void vol_read(const void *vol, int mode, size_t i, R *t, R *w)
{
    switch (mode) {
    case KFO_RG16I: {
        const int16_t *v = (const int16_t *)vol + 2 * i;
        *t = (R)v[0] * KC(0.00003051944088);
        *w = (R)v[1];
        break;
    }
    case KFO_F32X2: {
        const float *v = (const float *)vol + 2 * i;
        *t = (R)v[0]; *w = (R)v[1];
        break;
    }
    case KFO_F16X2: {
        const uint16_t *v = (const uint16_t *)vol + 2 * i;
        *t = (R)kfo_h2f(v[0]); *w = (R)kfo_h2f(v[1]);
        break;
    }
    default: {
        const double *v = (const double *)vol + 2 * i;
        *t = (R)v[0]; *w = (R)v[1];
    }
    }
}

/* set(): integrate.cs:45-50 -- RG16I stores int(v * 32766) (truncation) */
static inline void vol_write(void *vol, int mode, size_t i, R t, R w)
{
    switch (mode) {
    case KFO_RG16I: {
        int16_t *v = (int16_t *)vol + 2 * i;
        v[0] = (int16_t)(int)(t * KC(32766.0));
        v[1] = (int16_t)(int)w;
        break;
    }
    case KFO_F32X2: {
        float *v = (float *)vol + 2 * i;
        v[0] = (float)t; v[1] = (float)w;
        break;
    }
    case KFO_F16X2: {
        uint16_t *v = (uint16_t *)vol + 2 * i;
        v[0] = kfo_f2h((float)t); v[1] = kfo_f2h((float)w);
        break;
    }
    default: {
        double *v = (double *)vol + 2 * i;
        v[0] = (double)t; v[1] = (double)w;
    }
    }
}

/* raw value used by raycast.cs vs() (:59-67): the int16 itself for RG16I,
 * the stored value otherwise; zero for out-of-bounds loads. */
static inline R vol_raw(const void *vol, int mode, const int N[3], int x, int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= N[0] || y >= N[1] || z >= N[2]) return 0;
    size_t i = ((size_t)z * N[1] + y) * N[0] + x;
    switch (mode) {
    case KFO_RG16I: return (R)((const int16_t *)vol)[2 * i];
    case KFO_F32X2: return (R)((const float *)vol)[2 * i];
    case KFO_F16X2: return (R)kfo_h2f(((const uint16_t *)vol)[2 * i]);
    default: return (R)((const double *)vol)[2 * i];
    }
}

static inline R raw_scale(int mode)
{
    return mode == KFO_RG16I ? KC(0.00003051944088) : KC(1.0);
}

KFO_EXPORT void kfo_reset_volume(void *vol, int mode, const int N[3])
{
    memset(vol, 0, (size_t)N[0] * N[1] * N[2] * (size_t)kfo_voxel_bytes(mode));
}

/* ------------------------------------------------------------------------- */
/* integrate: gFusion.cpp:1234-1287 + integrate.cs:52-169                      */
/* ------------------------------------------------------------------------- */
KFO_EXPORT void kfo_integrate(void *vol, int mode, const int N[3], const R dim[3],
                              const R *depth, int dw, int dh,
                              const R pose[16], const R cam[4], R mu, R maxWeight)
{
    R inv[16], K[16];
    kfo_inverse4(pose, inv);
    kfo_camera_matrix(cam, K);
    R vz = dim[2] / (R)N[2];
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < N[1]; ++y)
        for (int x = 0; x < N[0]; ++x) {
            /* getVolumePosition(pix) at z = 0 (integrate.cs:16-19) */
            R c0 = ((R)x + KC(0.5)) * dim[0] / (R)N[0];
            R c1 = ((R)y + KC(0.5)) * dim[1] / (R)N[1];
            R c2 = ((R)0 + KC(0.5)) * dim[2] / (R)N[2];
            R pos[3], cx[3], dl[3], cd[3];
            opmul3(inv, c0, c1, c2, pos);
            opmul3(K, pos[0], pos[1], pos[2], cx);
            rotate3(inv, 0, 0, vz, dl);
            rotate3(K, dl[0], dl[1], dl[2], cd);
            for (int z = 0; z < N[2]; ++z,
                     pos[0] += dl[0], pos[1] += dl[1], pos[2] += dl[2],
                     cx[0] += cd[0], cx[1] += cd[1], cx[2] += cd[2]) {
                if (pos[2] < KC(0.0001)) continue;
                R px = cx[0] / cx[2] + KC(0.5);
                R py = cx[1] / cx[2] + KC(0.5);
                if (px < 0 || px > (R)(dw - 1) || py < 0 || py > (R)(dh - 1)) continue;
                int ix = (int)px, iy = (int)py;
                R d = depth[iy * dw + ix] / KC(1000.0);
                if (d == 0) continue;
                R qx = pos[0] / pos[2], qy = pos[1] / pos[2];
                R diff = (d - cx[2]) * KSQRT((KC(1.0) + qx * qx) + qy * qy);
                size_t vi = ((size_t)z * N[1] + y) * N[0] + x;
                if (diff > -mu) {
                    R sdf = gmin(KC(1.0), diff / mu);
                    R t, w;
                    vol_read(vol, mode, vi, &t, &w);
                    R wd = (w * t + sdf) / (w + KC(1.0));
                    if (wd < KC(0.2)) {
                        t = gmin(gmax(wd, KC(-0.1)), KC(0.2));
                        w = gmin(w + KC(1.0), maxWeight);
                    } else {
                        t = 0; w = 0;
                    }
                    vol_write(vol, mode, vi, t, w);
                } else {
                    vol_write(vol, mode, vi, 0, 0);
                }
            }
        }
}

/* ------------------------------------------------------------------------- */
/* raycast: gFusion.cpp:1289-1366 + raycast.cs:69-82, 133-276, 278-321         */
/* ------------------------------------------------------------------------- */
static R interp_vol(const void *vol, int mode, const int N[3], const R dim[3], R p0, R p1, R p2)
{
    R s[3] = {(p0 * (R)N[0] / dim[0]) - KC(0.5), (p1 * (R)N[1] / dim[1]) - KC(0.5),
              (p2 * (R)N[2] / dim[2]) - KC(0.5)};
    int lo[3], hi[3];
    R f[3];
    for (int k = 0; k < 3; ++k) {
        R fl = KFLOOR(s[k]);
        int b = f2i(fl);
        f[k] = s[k] - fl;
        lo[k] = (b == INT32_MIN) ? b : imax(b, 0);
        hi[k] = (b == INT32_MIN) ? b : imin(b + 1, N[k] - 1);
    }
#define V(a, b, c) vol_raw(vol, mode, N, a, b, c)
    R r = ((V(lo[0], lo[1], lo[2]) * (1 - f[0]) + V(hi[0], lo[1], lo[2]) * f[0]) * (1 - f[1]) +
           (V(lo[0], hi[1], lo[2]) * (1 - f[0]) + V(hi[0], hi[1], lo[2]) * f[0]) * f[1]) * (1 - f[2]) +
          ((V(lo[0], lo[1], hi[2]) * (1 - f[0]) + V(hi[0], lo[1], hi[2]) * f[0]) * (1 - f[1]) +
           (V(lo[0], hi[1], hi[2]) * (1 - f[0]) + V(hi[0], hi[1], hi[2]) * f[0]) * f[1]) * f[2];
#undef V
    return mode == KFO_RG16I ? r * KC(0.00003051944088) : r;
}

static void gradient(const void *vol, int mode, const int N[3], const R dim[3], const R hit[3], R g[3])
{
    R s[3], f[3];
    int ll[3], lu[3], ul[3], uu[3];
    for (int k = 0; k < 3; ++k) {
        s[k] = (hit[k] * (R)N[k] / dim[k]) - KC(0.5);
        R fl = KFLOOR(s[k]);
        int b = f2i(fl);
        f[k] = s[k] - fl;
        if (b == INT32_MIN) {
            ll[k] = lu[k] = ul[k] = uu[k] = b;
        } else {
            ll[k] = imax(b - 1, 0);
            lu[k] = imax(b, 0);
            ul[k] = imin(b + 1, N[k] - 1);
            uu[k] = imin(b + 2, N[k] - 1);
        }
    }
    const int *lo = lu, *hi = ul;
    R fx = f[0], fy = f[1], fz = f[2];
#define V(a, b, c) vol_raw(vol, mode, N, a, b, c)
    g[0] = (((V(ul[0], lo[1], lo[2]) - V(ll[0], lo[1], lo[2])) * (1 - fx) +
             (V(uu[0], lo[1], lo[2]) - V(lu[0], lo[1], lo[2])) * fx) * (1 - fy) +
            ((V(ul[0], hi[1], lo[2]) - V(ll[0], hi[1], lo[2])) * (1 - fx) +
             (V(uu[0], hi[1], lo[2]) - V(lu[0], hi[1], lo[2])) * fx) * fy) * (1 - fz) +
           (((V(ul[0], lo[1], hi[2]) - V(ll[0], lo[1], hi[2])) * (1 - fx) +
             (V(uu[0], lo[1], hi[2]) - V(lu[0], lo[1], hi[2])) * fx) * (1 - fy) +
            ((V(ul[0], hi[1], hi[2]) - V(ll[0], hi[1], hi[2])) * (1 - fx) +
             (V(uu[0], hi[1], hi[2]) - V(lu[0], hi[1], hi[2])) * fx) * fy) * fz;
    g[1] = (((V(lo[0], ul[1], lo[2]) - V(lo[0], ll[1], lo[2])) * (1 - fx) +
             (V(hi[0], ul[1], lo[2]) - V(hi[0], ll[1], lo[2])) * fx) * (1 - fy) +
            ((V(lo[0], uu[1], lo[2]) - V(lo[0], lu[1], lo[2])) * (1 - fx) +
             (V(hi[0], uu[1], lo[2]) - V(hi[0], lu[1], lo[2])) * fx) * fy) * (1 - fz) +
           (((V(lo[0], ul[1], hi[2]) - V(lo[0], ll[1], hi[2])) * (1 - fx) +
             (V(hi[0], ul[1], hi[2]) - V(hi[0], ll[1], hi[2])) * fx) * (1 - fy) +
            ((V(lo[0], uu[1], hi[2]) - V(lo[0], lu[1], hi[2])) * (1 - fx) +
             (V(hi[0], uu[1], hi[2]) - V(hi[0], lu[1], hi[2])) * fx) * fy) * fz;
    g[2] = (((V(lo[0], lo[1], ul[2]) - V(lo[0], lo[1], ll[2])) * (1 - fx) +
             (V(hi[0], lo[1], ul[2]) - V(hi[0], lo[1], ll[2])) * fx) * (1 - fy) +
            ((V(lo[0], hi[1], ul[2]) - V(lo[0], hi[1], ll[2])) * (1 - fx) +
             (V(hi[0], hi[1], ul[2]) - V(hi[0], hi[1], ll[2])) * fx) * fy) * (1 - fz) +
           (((V(lo[0], lo[1], uu[2]) - V(lo[0], lo[1], lu[2])) * (1 - fx) +
             (V(hi[0], lo[1], uu[2]) - V(hi[0], lo[1], lu[2])) * fx) * (1 - fy) +
            ((V(lo[0], hi[1], uu[2]) - V(lo[0], hi[1], lu[2])) * (1 - fx) +
             (V(hi[0], hi[1], uu[2]) - V(hi[0], hi[1], lu[2])) * fx) * fy) * fz;
#undef V
    R gs = KC(0.5) * raw_scale(mode);
    for (int k = 0; k < 3; ++k) g[k] = (g[k] * (dim[k] / (R)N[k])) * gs;
}

/* raycast.cs:133-225 for one pixel; returns 1 on a hit (hit = xyz, t) */
static int march(const void *vol, int mode, const int N[3], const R dim[3], const R view[16],
                 R nearPlane, R farPlane, R step, R largeStep, int x, int y, R hit[4])
{
    R o[3] = {view[12], view[13], view[14]};
    R d[3];
    rotate3(view, (R)x, (R)y, KC(1.0), d);
    R tmin[3], tmax[3];
    for (int k = 0; k < 3; ++k) {
        R invR = KC(1.0) / d[k];
        R tbot = (KC(-1.0) * invR) * o[k];
        R ttop = invR * (dim[k] - o[k]);
        tmin[k] = gmin(ttop, tbot);
        tmax[k] = gmax(ttop, tbot);
    }
    R largest_tmin = gmax(gmax(tmin[0], tmin[1]), gmax(tmin[0], tmin[2]));
    R smallest_tmax = gmin(gmin(tmax[0], tmax[1]), gmin(tmax[0], tmax[2]));
    R tnear = gmax(largest_tmin, nearPlane);
    R tfar = gmin(smallest_tmax, farPlane);
    if (!(tnear < tfar)) return 0;
    R t = tnear, stepsize = largeStep;
    R f_t = interp_vol(vol, mode, N, dim, o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t);
    R f_tt = 0;
    if (f_t >= 0) {
        for (; t < tfar; t += stepsize) {
            f_tt = interp_vol(vol, mode, N, dim, o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t);
            if (f_tt < 0) break;
            if (f_tt < KC(0.8)) stepsize = step;
            f_t = f_tt;
        }
        if (f_tt < 0) {
            t = t + stepsize * f_tt / (f_t - f_tt);
            hit[0] = o[0] + d[0] * t;
            hit[1] = o[1] + d[1] * t;
            hit[2] = o[2] + d[2] * t;
            hit[3] = t;
            return 1;
        }
    }
    return 0;
}

/* view = m_pose * getInverseCameraMatrix(camPams)   (gFusion.cpp:1297-1298) */
KFO_EXPORT void kfo_raycast_view(const R pose[16], const R cam[4], R view[16])
{
    R iK[16];
    kfo_inverse_camera_matrix(cam, iK);
    kfo_mul4(pose, iK, view);
}

/* raycast.cs main (:278-321) over a (w x h) image */
KFO_EXPORT void kfo_raycast(const void *vol, int mode, const int N[3], const R dim[3],
                            const R pose[16], const R cam[4], R nearPlane, R farPlane,
                            R step, R largeStep, int w, int h, R *refV, R *refN)
{
    R view[16];
    kfo_raycast_view(pose, cam, view);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            R hit[4];
            R *ov = refV + 4 * (y * w + x), *on = refN + 4 * (y * w + x);
            if (march(vol, mode, N, dim, view, nearPlane, farPlane, step, largeStep, x, y, hit) &&
                hit[3] > 0) {
                ov[0] = hit[0]; ov[1] = hit[1]; ov[2] = hit[2]; ov[3] = KC(1.0);
                R g[3];
                gradient(vol, mode, N, dim, hit, g);
                R len = KSQRT(dot3(g[0], g[1], g[2], g[0], g[1], g[2]));
                if (len == 0) {
                    on[0] = on[1] = on[2] = on[3] = 0;
                } else {
                    on[0] = g[0] / len; on[1] = g[1] / len; on[2] = g[2] / len; on[3] = KC(1.0);
                }
            } else {
                ov[0] = ov[1] = ov[2] = ov[3] = 0;
                on[0] = on[1] = on[2] = on[3] = 0;
            }
        }
}

/* ------------------------------------------------------------------------- */
/* intensityProjection: gFusion.cpp:1368-1409 + mip.cs:33-107                  */
/* ------------------------------------------------------------------------- */
KFO_EXPORT void kfo_intensity_projection(const void *vol, int mode, const int N[3], const R dim[3],
                                         const R pose[16], const R cam[4], R nearPlane, R farPlane,
                                         R largeStep, int w, int h, R *out)
{
    R view[16];
    kfo_raycast_view(pose, cam, view);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            R o[3] = {view[12], view[13], view[14]};
            R d[3];
            rotate3(view, (R)x, (R)y, KC(1.0), d);
            R tmin[3], tmax[3];
            for (int k = 0; k < 3; ++k) {
                R invR = KC(1.0) / d[k];
                R tbot = (KC(-1.0) * invR) * o[k];
                R ttop = invR * (dim[k] - o[k]);
                tmin[k] = gmin(ttop, tbot);
                tmax[k] = gmax(ttop, tbot);
            }
            R largest_tmin = gmax(gmax(tmin[0], tmin[1]), gmax(tmin[0], tmin[2]));
            R smallest_tmax = gmin(gmin(tmax[0], tmax[1]), gmin(tmax[0], tmax[2]));
            R tnear = gmax(largest_tmin, nearPlane);
            R tfar = gmin(smallest_tmax, farPlane);
            R sum = 0, zsum = 0;
            if (tnear < tfar) {
                for (R t = tnear; t < tfar; t += largeStep) {
                    R f = interp_vol(vol, mode, N, dim, o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t);
                    if (f > 0) sum += f;
                    else if (f == 0) zsum += f;
                }
            }
            R *px = out + 4 * ((size_t)y * w + x);
            px[0] = sum; px[1] = 0; px[2] = zsum; px[3] = KC(1.0);
        }
}

/* ------------------------------------------------------------------------- */
/* Point-to-volume tracking: gFusion.cpp:835-1038 (TrackSDF), :1040-1084      */
/* (trackSDF), :1086-1112 (reduceSDF), :1195-1232 (getSDFReduction),          */
/* trackSDF.cs:186-220, 222-235, 296-306, 442-530, reduceSDF.cs:28-116.       */
/* ------------------------------------------------------------------------- */

/* imageLoad(volumeData, p).x * 0.00003051944088f (trackSDF.cs:195-205): the
 * stored value in truncation units, zero out of bounds */
static inline R sdf_voxel(const void *vol, int mode, const int N[3], int x, int y, int z)
{
    return vol_raw(vol, mode, N, x, y, z) * raw_scale(mode);
}

/* float(J) + 1.0 converted back to int (ivec3(I, J + 1.0, K)) */
static inline int ip1(int v) { return v == INT32_MIN ? v : v + 1; }

/* SDF(location), trackSDF.cs:186-220: modf splits each coordinate into an
 * integer part (toward zero) and a signed fraction; corners (I..I+1, J..J+1,
 * K..K+1) are blended z first, then y, then x. */
KFO_EXPORT R kfo_sdf_sample(const void *vol, int mode, const int N[3], R lx, R ly, R lz)
{
    R i, j, k;
    R x = KMODF(lx, &i), y = KMODF(ly, &j), z = KMODF(lz, &k);
    int I = f2i(i), J = f2i(j), K = f2i(k);
    R N0 = sdf_voxel(vol, mode, N, I, J, K);
    R N1 = sdf_voxel(vol, mode, N, I, ip1(J), K);
    R N2 = sdf_voxel(vol, mode, N, ip1(I), J, K);
    R N3 = sdf_voxel(vol, mode, N, ip1(I), ip1(J), K);
    R N4 = sdf_voxel(vol, mode, N, I, J, ip1(K));
    R N5 = sdf_voxel(vol, mode, N, I, ip1(J), ip1(K));
    R N6 = sdf_voxel(vol, mode, N, ip1(I), J, ip1(K));
    R N7 = sdf_voxel(vol, mode, N, ip1(I), ip1(J), ip1(K));
    R a0 = N0 * (KC(1.0) - z) + N4 * z;
    R a1 = N1 * (KC(1.0) - z) + N5 * z;
    R b0 = N2 * (KC(1.0) - z) + N6 * z;
    R b1 = N3 * (KC(1.0) - z) + N7 * z;
    return (a0 * (KC(1.0) - y) + a1 * y) * (KC(1.0) - x) + (b0 * (KC(1.0) - y) + b1 * y) * x;
}

/* trackSDF.cs:460-530 for one pixel: rec = {h, D, J[6]} */
KFO_EXPORT void kfo_track_sdf_pixel(const void *vol, int mode, const int N[3], const R dim[3], const R T[16],
                                    const R v[3], R rec[8])
{
    R cp[4] = {v[0], v[1], v[2], KC(1.0)};
    R pv[4];
    matvec4(T, cp, pv); /* projectedVertex = Ttrack * vec4(cameraPoint.xyz, 1) */
    /* cvp (:474): x and y lose 0.5 twice (inside and outside the vec3), z once */
    R c0 = ((pv[0] * (R)N[0] / dim[0]) - KC(0.5)) - KC(0.5);
    R c1 = ((pv[1] * (R)N[1] / dim[1]) - KC(0.5)) - KC(0.5);
    R c2 = (pv[2] * (R)N[2] / dim[2]) - KC(0.5);
    R D = kfo_sdf_sample(vol, mode, N, c0, c1, c2);
    /* SDFGradient (:223-235): central difference over +-1 voxel index, divided
     * by 2 * (volDim.x / volSize.x) for every axis */
    R den = KC(2.0) * (dim[0] / (R)N[0]);
    R gx = (kfo_sdf_sample(vol, mode, N, c0 + KC(1.0), c1 + KC(0.0), c2 + KC(0.0)) -
            kfo_sdf_sample(vol, mode, N, c0 - KC(1.0), c1 - KC(0.0), c2 - KC(0.0))) / den;
    R gy = (kfo_sdf_sample(vol, mode, N, c0 + KC(0.0), c1 + KC(1.0), c2 + KC(0.0)) -
            kfo_sdf_sample(vol, mode, N, c0 - KC(0.0), c1 - KC(1.0), c2 - KC(0.0))) / den;
    R gz = (kfo_sdf_sample(vol, mode, N, c0 + KC(0.0), c1 + KC(0.0), c2 + KC(1.0)) -
            kfo_sdf_sample(vol, mode, N, c0 - KC(0.0), c1 - KC(0.0), c2 - KC(1.0))) / den;
    R absD = D < 0 ? -D : D;
    for (int k = 0; k < 8; ++k) rec[k] = 0;
    if (D < KC(1.0) && D > KC(-1.0)) {
        /* getJ(dSDF_dx, dx_dxi) (:296-306) with the dx_dxi of :492-497 */
        R px = pv[0], py = pv[1], pz = pv[2];
        rec[0] = absD < KC(0.02) ? KC(1.0) : KC(0.02) / absD; /* huber (:522) */
        rec[1] = D;
        rec[2] = (gx * KC(0.0) + gy * (-pz)) + gz * py;
        rec[3] = (gx * pz + gy * KC(0.0)) + gz * (-px);
        rec[4] = (gx * (-py) + gy * px) + gz * KC(0.0);
        rec[5] = (gx * KC(1.0) + gy * KC(0.0)) + gz * KC(0.0);
        rec[6] = (gx * KC(0.0) + gy * KC(1.0)) + gz * KC(0.0);
        rec[7] = (gx * KC(0.0) + gy * KC(0.0)) + gz * KC(1.0);
    }
}

/* trackSDF.cs main over a (w x h) level: writes the 8-value record of every
 * pixel with x < w - 1 (the shader's bounds test, :455) into buf[y*w + x];
 * records of the last column keep whatever an earlier dispatch left there. */
KFO_EXPORT void kfo_track_sdf(const void *vol, int mode, const int N[3], const R dim[3], const R *vert, int w,
                              int h, const R T[16], R *buf)
{
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w - 1; ++x) {
            const R *v = vert + 4 * (y * w + x);
            kfo_track_sdf_pixel(vol, mode, N, dim, T, v, buf + 8 * (y * w + x));
        }
}

/* reduceSDF.cs:28-116, dispatched as 8 WGs x 112 threads (gFusion.cpp:1099):
 * serial FP32 sums per thread in row order, D^2 added twice (:52, :60), the
 * count [28] over every pixel of the rows, then the 112 rows added serially. */
KFO_EXPORT void kfo_reduce_sdf(const R *buf, int w, int h, R out[256])
{
    R (*S)[112][32] = (R (*)[112][32])malloc(sizeof(R) * 8 * 112 * 32);
#pragma omp parallel for schedule(static)
    for (int gs = 0; gs < 8 * 112; ++gs) {
        int g = gs / 112, s = gs % 112;
        R sums[32];
        for (int k = 0; k < 32; ++k) sums[k] = 0;
        for (int y = g; y < h; y += 8)
            for (int x = s; x < w; x += 112) {
                const R *r = buf + 8 * (y * w + x);
                R hh = r[0], D = r[1];
                const R *J = r + 2;
                sums[0] += D * D;
                for (int i = 0; i < 6; ++i) sums[i + 1] += hh * J[i] * D;
                sums[0] += D * D;
                int q = 7;
                for (int a = 0; a < 6; ++a)
                    for (int b = a; b < 6; ++b) sums[q++] += hh * J[a] * J[b];
                sums[28] += KC(1.0);
            }
        memcpy(S[g][s], sums, sizeof(sums));
    }
    for (int g = 0; g < 8; ++g)
        for (int k = 0; k < 32; ++k) {
            R acc = S[g][0][k];
            for (int s = 1; s < 112; ++s) acc += S[g][s][k];
            out[g * 32 + k] = acc;
        }
    free(S);
}

/* Eigen::LDLT<MatrixXd>(A).solve(b) (gFusion.cpp:918): in-place LDL^T of the
 * lower triangle with symmetric pivoting on the largest remaining |diagonal|,
 * then P^T L^-T D^+ L^-1 P b where D^+ zeroes pivots with |d| <= DBL_MIN
 * (Eigen's LDLT::_solve_impl).  Inner products are summed in index order. */
KFO_EXPORT void kfo_ldlt_solve6(const double Ain[36], const double b[6], double x[6])
{
    double m[6][6];
    int tr[6];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) m[i][j] = Ain[i * 6 + j];
    for (int k = 0; k < 6; ++k) {
        int big = k;
        double bv = fabs(m[k][k]);
        for (int i = k + 1; i < 6; ++i)
            if (fabs(m[i][i]) > bv) { bv = fabs(m[i][i]); big = i; }
        tr[k] = big;
        if (big != k) {
            for (int j = 0; j < k; ++j) { double t = m[k][j]; m[k][j] = m[big][j]; m[big][j] = t; }
            for (int i = big + 1; i < 6; ++i) { double t = m[i][k]; m[i][k] = m[i][big]; m[i][big] = t; }
            { double t = m[k][k]; m[k][k] = m[big][big]; m[big][big] = t; }
            for (int i = k + 1; i < big; ++i) { double t = m[i][k]; m[i][k] = m[big][i]; m[big][i] = t; }
        }
        double temp[6];
        if (k > 0) {
            for (int j = 0; j < k; ++j) temp[j] = m[j][j] * m[k][j];
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += m[k][j] * temp[j];
            m[k][k] -= s;
            for (int i = k + 1; i < 6; ++i) {
                double t = 0.0;
                for (int j = 0; j < k; ++j) t += m[i][j] * temp[j];
                m[i][k] -= t;
            }
        }
        double akk = m[k][k];
        int valid = fabs(akk) > 0.0;
        if (k == 0 && !valid) {
            /* the whole diagonal is zero (or NaN): Eigen stops here with the
             * identity permutation and the matrix left as it is; the solve
             * below still runs on it */
            for (int j = 0; j < 6; ++j) tr[j] = j;
            break;
        }
        if (valid)
            for (int i = k + 1; i < 6; ++i) m[i][k] /= akk;
    }
    double y[6];
    for (int i = 0; i < 6; ++i) y[i] = b[i];
    for (int k = 0; k < 6; ++k) { double t = y[k]; y[k] = y[tr[k]]; y[tr[k]] = t; }
    for (int i = 0; i < 6; ++i) { /* unit lower solve */
        double s = y[i];
        for (int j = 0; j < i; ++j) s -= m[i][j] * y[j];
        y[i] = s;
    }
    for (int i = 0; i < 6; ++i) y[i] = (fabs(m[i][i]) > DBL_MIN) ? y[i] / m[i][i] : 0.0;
    for (int i = 5; i >= 0; --i) { /* unit upper (L^T) solve */
        double s = y[i];
        for (int j = i + 1; j < 6; ++j) s -= m[j][i] * y[j];
        y[i] = s;
    }
    for (int k = 5; k >= 0; --k) { double t = y[k]; y[k] = y[tr[k]]; y[tr[k]] = t; }
    for (int i = 0; i < 6; ++i) x[i] = y[i];
}

/* Twist(xi).exp() (gFusion.cpp:810-833, :866): the 4x4 twist matrix holds xi
 * as R ([w]x = rows (0,-w2,w1), (w2,0,-w0), (-w1,w0,0), translation xi3..5);
 * its matrix exponential in closed form (Rodrigues rotation, V = I + B[w]x +
 * C[w]x^2), evaluated in double and rounded to R.  Column-major out. */
KFO_EXPORT void kfo_twist_exp(const double xi[6], R out[16])
{
    double w0 = (double)(R)xi[0], w1 = (double)(R)xi[1], w2 = (double)(R)xi[2];
    double v[3] = {(double)(R)xi[3], (double)(R)xi[4], (double)(R)xi[5]};
    double th2 = w0 * w0 + w1 * w1 + w2 * w2;
    double A, B, Cc;
    if (th2 < 1e-8) {
        A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
        Cc = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
    } else {
        double th = sqrt(th2), s = sin(th), c = cos(th);
        A = s / th;
        B = (1.0 - c) / th2;
        Cc = (th - s) / (th2 * th);
    }
    double W[3][3] = {{0.0, -w2, w1}, {w2, 0.0, -w0}, {-w1, w0, 0.0}};
    double W2[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) W2[r][c] = W[r][0] * W[0][c] + W[r][1] * W[1][c] + W[r][2] * W[2][c];
    double Rm[3][3], V[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double I = (r == c) ? 1.0 : 0.0;
            Rm[r][c] = I + A * W[r][c] + B * W2[r][c];
            V[r][c] = I + B * W[r][c] + Cc * W2[r][c];
        }
    for (int c = 0; c < 3; ++c) {
        for (int r = 0; r < 3; ++r) out[c * 4 + r] = (R)Rm[r][c];
        out[c * 4 + 3] = 0;
    }
    for (int r = 0; r < 3; ++r) out[12 + r] = (R)(V[r][0] * v[0] + V[r][1] * v[1] + V[r][2] * v[2]);
    out[15] = KC(1.0);
}

/* One Gauss-Newton step of TrackSDF (gFusion.cpp:878-931): scale C and b by
 * 1 / max(C), damp with 0.01 * iteration on the diagonal, result -= LDLT
 * solve; returns |result - result_prev| and updates both. */
KFO_EXPORT double kfo_sdf_step(const R b[6], const R C[36], int iteration, double result[6], double result_prev[6])
{
    double dC[36], db[6], dx[6];
    double mx = (double)C[0];
    for (int k = 0; k < 36; ++k) {
        dC[k] = (double)C[k];
        if (dC[k] > mx) mx = dC[k];
    }
    double scaling = 1.0 / mx;
    for (int k = 0; k < 36; ++k) dC[k] *= scaling;
    for (int k = 0; k < 6; ++k) db[k] = (double)b[k] * scaling;
    for (int k = 0; k < 6; ++k) dC[k * 6 + k] += (double)iteration * 0.01;
    kfo_ldlt_solve6(dC, db, dx);
    double cn = 0.0;
    for (int k = 0; k < 6; ++k) {
        result[k] = result[k] - dx[k];
        double d = result[k] - result_prev[k];
        cn += d * d;
    }
    for (int k = 0; k < 6; ++k) result_prev[k] = result[k];
    return sqrt(cn);
}

/* ------------------------------------------------------------------------- */
/* Surface export: MCubes::histoPyramids (mcubes.cpp:206-390), exportMesh     */
/* (:429-509), histoPyramids.cs (classifyCubes :104-170, constructHPLevel      */
/* :42-90), traverseHistoPyramids.cs (scanHPLevel :52-109, traverseHPLevel     */
/* :126-256).                                                                  */
/* ------------------------------------------------------------------------- */

/* Cube corners in the order of the classification bits (histoPyramids.cs:
 * 156-163 reads cubeOffsets 0,1,3,2,4,5,7,6 into bits 0..7) and the 12 edges
 * as (corner, corner); edge k joins the corners listed, in that order. */
static const int MC_CORNER[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                                    {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
static const int MC_EDGE[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                   {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
static const int MC_FACE[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                  {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};

static int mc_edge_of(int a, int b)
{
    for (int k = 0; k < 12; ++k)
        if ((MC_EDGE[k][0] == a && MC_EDGE[k][1] == b) || (MC_EDGE[k][0] == b && MC_EDGE[k][1] == a)) return k;
    return -1;
}

/* Triangle table built from the cube's topology: the surface crosses every
 * edge whose corners lie on different sides of the iso level (bit set = below
 * it); on each face the crossings are joined in pairs, a face with four
 * crossings separating its two below-iso corners; the joined segments close
 * into polygons, each oriented with its normal towards the above-iso side and
 * split into a fan of triangles from its lowest edge.  tri[c][0..3n) holds the
 * n triangles of case c as edge indices, 255 after them; ntri[c] = n. */
KFO_EXPORT void kfo_mc_tables(uint8_t tri[256][16], uint8_t ntri[256])
{
    for (int c = 0; c < 256; ++c) {
        int in[8];
        for (int v = 0; v < 8; ++v) in[v] = (c >> v) & 1;
        int nb[12][2], deg[12];
        for (int e = 0; e < 12; ++e) deg[e] = 0;
        for (int f = 0; f < 6; ++f) {
            int es[4], cr[4], n = 0;
            for (int i = 0; i < 4; ++i) {
                int a = MC_FACE[f][i], b = MC_FACE[f][(i + 1) % 4];
                es[i] = mc_edge_of(a, b);
                cr[i] = in[a] != in[b];
                n += cr[i];
            }
            int seg[2][2], ns = 0;
            if (n == 2) {
                int k = 0;
                for (int i = 0; i < 4; ++i)
                    if (cr[i]) seg[0][k++] = es[i];
                ns = 1;
            } else if (n == 4) {
                if (in[MC_FACE[f][0]]) { /* corners 0 and 2 are below: cut them off */
                    seg[0][0] = es[3]; seg[0][1] = es[0];
                    seg[1][0] = es[1]; seg[1][1] = es[2];
                } else {
                    seg[0][0] = es[0]; seg[0][1] = es[1];
                    seg[1][0] = es[2]; seg[1][1] = es[3];
                }
                ns = 2;
            }
            for (int k = 0; k < ns; ++k) {
                int a = seg[k][0], b = seg[k][1];
                nb[a][deg[a]++] = b;
                nb[b][deg[b]++] = a;
            }
        }
        for (int i = 0; i < 16; ++i) tri[c][i] = 255;
        int used[12] = {0}, nt = 0;
        for (int e0 = 0; e0 < 12; ++e0) {
            if (!deg[e0] || used[e0]) continue;
            int cyc[12], m = 0, prev = -1, cur = e0;
            for (;;) {
                cyc[m++] = cur;
                used[cur] = 1;
                int nx = nb[cur][0] != prev ? nb[cur][0] : nb[cur][1];
                if (nx == e0 || used[nx]) break;
                prev = cur;
                cur = nx;
            }
            /* orientation: Newell normal of the edge midpoints against the
             * below -> above directions of the polygon's edges */
            double P[12][3], nrm[3] = {0, 0, 0};
            for (int i = 0; i < m; ++i)
                for (int k = 0; k < 3; ++k)
                    P[i][k] = 0.5 * (MC_CORNER[MC_EDGE[cyc[i]][0]][k] + MC_CORNER[MC_EDGE[cyc[i]][1]][k]);
            for (int i = 0; i < m; ++i) {
                const double *a = P[i], *b = P[(i + 1) % m];
                nrm[0] += (a[1] - b[1]) * (a[2] + b[2]);
                nrm[1] += (a[2] - b[2]) * (a[0] + b[0]);
                nrm[2] += (a[0] - b[0]) * (a[1] + b[1]);
            }
            double dir = 0;
            for (int i = 0; i < m; ++i) {
                int va = MC_EDGE[cyc[i]][0], vb = MC_EDGE[cyc[i]][1];
                int lo = in[va] ? va : vb, hi = in[va] ? vb : va;
                for (int k = 0; k < 3; ++k) dir += nrm[k] * (MC_CORNER[hi][k] - MC_CORNER[lo][k]);
            }
            if (dir < 0)
                for (int i = 1, j = m - 1; i < j; ++i, --j) { int t = cyc[i]; cyc[i] = cyc[j]; cyc[j] = t; }
            for (int i = 1; i + 1 < m; ++i) {
                tri[c][3 * nt + 0] = (uint8_t)cyc[0];
                tri[c][3 * nt + 1] = (uint8_t)cyc[i];
                tri[c][3 * nt + 2] = (uint8_t)cyc[i + 1];
                ++nt;
            }
        }
        ntri[c] = (uint8_t)nt;
    }
}

/* stored value at a voxel as histoPyramids.cs getVolumeData reads it for the
 * RG16I volume (float of the int16), the value itself for the other modes */
static inline R mc_field(const void *vol, int mode, const int N[3], int x, int y, int z)
{
    return vol_raw(vol, mode, N, x, y, z);
}

/* classifyCubes (histoPyramids.cs:104-170) for every voxel: (numTri, cubeIndex);
 * the 10-voxel border keeps (0, 0) (the texture's initial contents). */
KFO_EXPORT void kfo_mc_classify(const void *vol, int mode, const int N[3], R iso, uint8_t *ntri_out,
                                uint8_t *cube_out)
{
    uint8_t tri[256][16], ntri[256];
    kfo_mc_tables(tri, ntri);
    static const int order[8] = {0, 1, 3, 2, 4, 5, 7, 6}; /* cubeOffsets index of bit k */
    static const int OFF[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1},
                                  {0, 1, 0}, {1, 1, 0}, {0, 1, 1}, {1, 1, 1}};
#pragma omp parallel for schedule(static)
    for (int z = 0; z < N[2]; ++z)
        for (int y = 0; y < N[1]; ++y)
            for (int x = 0; x < N[0]; ++x) {
                size_t i = ((size_t)z * N[1] + y) * N[0] + x;
                ntri_out[i] = 0;
                cube_out[i] = 0;
                if (x <= 10 || x >= N[0] - 10 || y <= 10 || y >= N[1] - 10 || z <= 10 || z >= N[2] - 10)
                    continue;
                R field[8];
                for (int k = 0; k < 8; ++k)
                    field[k] = mc_field(vol, mode, N, x + OFF[k][0], y + OFF[k][1], z + OFF[k][2]);
                int pos = 0;
                for (int k = 0; k < 7; ++k) /* field[7] is not looked at (:143) */
                    if (field[k] > 0) pos = 1;
                if (!pos) continue;
                int ci = 0;
                for (int k = 0; k < 8; ++k) ci |= (field[order[k]] < iso) << k;
                ntri_out[i] = ntri[ci];
                cube_out[i] = (uint8_t)ci;
            }
}

/* The whole histopyramid export for a cubic power-of-two volume: build the
 * levels by 2x2x2 sums (constructHPLevel), read the total at the top, then
 * for every triangle index walk the levels top-down (scanHPLevel) to its cube
 * and emit the triangle's three vertices (traverseHPLevel :190-250) as
 * (x, y, z, 0) in voxel-index units.  Returns the number of triangles; writes
 * at most `cap` triangles (3 x 4 values each) into `pos` (may be NULL). */
KFO_EXPORT long kfo_mc_extract(const void *vol, int mode, int n, R iso, R *pos, long cap)
{
    int N[3] = {n, n, n};
    int L = 0;
    while ((1 << L) < n) ++L;
    size_t nv = (size_t)n * n * n;
    uint8_t *nt0 = (uint8_t *)malloc(nv), *ci0 = (uint8_t *)malloc(nv);
    kfo_mc_classify(vol, mode, N, iso, nt0, ci0);
    uint32_t **lev = (uint32_t **)calloc((size_t)L + 1, sizeof(uint32_t *));
    for (int l = 1; l <= L; ++l) {
        int m = n >> l, pm = n >> (l - 1);
        lev[l] = (uint32_t *)calloc((size_t)m * m * m, 4);
        for (int z = 0; z < m; ++z)
            for (int y = 0; y < m; ++y)
                for (int x = 0; x < m; ++x) {
                    uint32_t acc = 0;
                    for (int k = 0; k < 8; ++k) {
                        int X = 2 * x + (k & 1), Z = 2 * z + ((k >> 1) & 1), Y = 2 * y + (k >> 2);
                        size_t j = ((size_t)Z * pm + Y) * pm + X;
                        acc += l == 1 ? nt0[j] : lev[l - 1][j];
                    }
                    lev[l][((size_t)z * m + y) * m + x] = acc;
                }
    }
    long total = L > 0 ? (long)lev[L][0] : (long)nt0[0];
    uint8_t tri[256][16], ntri[256];
    kfo_mc_tables(tri, ntri);
    long lim = pos ? (total < cap ? total : cap) : 0;
#pragma omp parallel for schedule(static)
    for (long t = 0; t < lim; ++t) {
        uint32_t target = (uint32_t)t;
        uint32_t cx = 0, cy = 0, cz = 0, cw = 0;
        for (int l = L - 1; l >= 0; --l) {
            int m = n >> l;
            uint32_t nb[8];
            for (int k = 0; k < 8; ++k) {
                uint32_t X = cx + (k & 1), Z = cz + ((k >> 1) & 1), Y = cy + (k >> 2);
                size_t j = ((size_t)Z * m + Y) * m + X;
                nb[k] = (X < (uint32_t)m && Y < (uint32_t)m && Z < (uint32_t)m) ? (l == 0 ? nt0[j] : lev[l][j]) : 0;
            }
            uint32_t acc = cw, child = 0, add = 0;
            for (int k = 0; k < 7; ++k) {
                acc += nb[k];
                if (acc <= target) { child++; add += nb[k]; }
            }
            cx += child & 1; cz += (child >> 1) & 1; cy += child >> 2;
            cx *= 2; cy *= 2; cz *= 2;
            cw += add;
        }
        cx /= 2; cy /= 2; cz /= 2;
        int ci = ci0[((size_t)cz * n + cy) * n + cx];
        for (int v = 0; v < 3; ++v) {
            int e = tri[ci][3 * (target - cw) + v];
            const int *a = MC_CORNER[MC_EDGE[e][0]], *b = MC_CORNER[MC_EDGE[e][1]];
            int p0[3] = {(int)cx + a[0], (int)cy + a[1], (int)cz + a[2]};
            int p1[3] = {(int)cx + b[0], (int)cy + b[1], (int)cz + b[2]};
            R v0 = mc_field(vol, mode, N, p0[0], p0[1], p0[2]);
            R diff = (iso - v0) / (mc_field(vol, mode, N, p1[0], p1[1], p1[2]) - v0);
            if (diff > 1 || diff < 0) diff = KC(0.5);
            R *o = pos + 4 * (3 * t + v);
            /* mix(p0, p1, diff) = p0 * (1 - diff) + p1 * diff */
            for (int k = 0; k < 3; ++k) o[k] = (R)p0[k] * (KC(1.0) - diff) + (R)p1[k] * diff;
            o[3] = 0;
        }
    }
    for (int l = 1; l <= L; ++l) free(lev[l]);
    free(lev); free(nt0); free(ci0);
    return total;
}

/* ------------------------------------------------------------------------- */
/* Session: gFusion host state + the frame block of kinect.cpp:654-695        */
/* ------------------------------------------------------------------------- */
#define KFO_LIB_MAX 256

typedef struct {
    int vol[3];
    R dim[3];
    R mu, maxWeight, nearPlane, farPlane, dist_threshold, normal_threshold;
    int iterations[3]; /* by level, 0 = finest */
    R cam[4];          /* fx, fy, cx, cy */
    int width, height; /* depth size (512 x 424) */
    int mode;          /* TSDF storage */
    int rule;          /* reduction combine rule */
} kfo_config;

typedef struct {
    kfo_config cfg;
    R *depth[3], *vert[3], *norm[3];
    R *refV, *refN;
    void *vol;
    R pose[16];
    int *result;
    R *error, *J;
    R partials[256];
    int counter, reset;
    R last_ae;
    int last_tracked;
    int iters_run;
    /* per-iteration log of the last Track(): level, AE, x[6] */
    int log_n;
    double log[64][8];
    /* point-to-volume tracking state (gFusion.h:426, :445) */
    R pose_eig[16];      /* m_pose_eig, column-major */
    double cum_twist[6]; /* m_cumTwist */
    R *sdf_buf;          /* m_bufferSDFReduction: 8 values per level-0 pixel */
    R sdf_partials[256];
    int sdf_log_n;       /* last TrackSDF(): level, iteration, AE, Cnorm, result[6] */
    double sdf_log[64][10];
    /* pose library (gFusion.h:424): poses only, the depth copies are never read */
    int lib_n;
    R lib[KFO_LIB_MAX][16];
    int track_mode;      /* 1 = point-to-plane (Track), 2 = point-to-volume (TrackSDF), 3 = both */
    int recover;         /* call recoverPose() when tracking fails (kinect.cpp:686-689) */
} kfo_session;

KFO_EXPORT int kfo_config_size(void) { return (int)sizeof(kfo_config); }

/* stepSize(): include/gFusion.h:72-81 */
static R step_size(const kfo_config *c)
{
    R minxyz = gmin(gmin(c->dim[0], c->dim[1]), c->dim[2]);
    R maxxyz = gmax(gmax((R)c->vol[0], (R)c->vol[1]), (R)c->vol[2]);
    return minxyz / maxxyz;
}

KFO_EXPORT kfo_session *kfo_session_create(const kfo_config *cfg)
{
    kfo_session *s = (kfo_session *)calloc(1, sizeof(kfo_session));
    s->cfg = *cfg;
    int w = cfg->width, h = cfg->height;
    for (int l = 0; l < 3; ++l) {
        size_t n = (size_t)(w >> l) * (h >> l);
        s->depth[l] = (R *)calloc(n, sizeof(R));
        s->vert[l] = (R *)calloc(4 * n, sizeof(R));
        s->norm[l] = (R *)calloc(4 * n, sizeof(R));
    }
    s->refV = (R *)calloc(4 * (size_t)w * h, sizeof(R));
    s->refN = (R *)calloc(4 * (size_t)w * h, sizeof(R));
    s->result = (int *)calloc((size_t)w * h, sizeof(int));
    s->error = (R *)calloc((size_t)w * h, sizeof(R));
    s->J = (R *)calloc(6 * (size_t)w * h, sizeof(R));
    s->vol = calloc((size_t)cfg->vol[0] * cfg->vol[1] * cfg->vol[2], (size_t)kfo_voxel_bytes(cfg->mode));
    /* glBufferData(..., NULL) at gFusion.cpp:349: contents taken as zero */
    s->sdf_buf = (R *)calloc(8 * (size_t)w * h, sizeof(R));
    for (int i = 0; i < 16; ++i) s->pose[i] = s->pose_eig[i] = (i % 5 == 0) ? KC(1.0) : KC(0.0);
    s->track_mode = 1;
    s->recover = 1;
    return s;
}

KFO_EXPORT void kfo_session_destroy(kfo_session *s)
{
    if (!s) return;
    for (int l = 0; l < 3; ++l) {
        free(s->depth[l]); free(s->vert[l]); free(s->norm[l]);
    }
    free(s->refV); free(s->refN); free(s->result); free(s->error); free(s->J); free(s->vol);
    free(s->sdf_buf);
    free(s);
}

KFO_EXPORT void kfo_session_set_pose(kfo_session *s, const R pose[16]) { memcpy(s->pose, pose, 16 * sizeof(R)); }
KFO_EXPORT void kfo_session_get_pose(const kfo_session *s, R pose[16]) { memcpy(pose, s->pose, 16 * sizeof(R)); }

KFO_EXPORT void *kfo_session_buffer(kfo_session *s, int which, int level)
{
    switch (which) {
    case 0: return s->depth[level];
    case 1: return s->vert[level];
    case 2: return s->norm[level];
    case 3: return s->refV;
    case 4: return s->refN;
    case 5: return s->vol;
    case 6: return s->result;
    case 7: return s->error;
    case 8: return s->J;
    case 9: return s->partials;
    case 10: return s->log;
    case 11: return s->sdf_buf;
    case 12: return s->sdf_partials;
    case 13: return s->sdf_log;
    case 14: return s->lib;
    case 15: return s->pose_eig;
    case 16: return s->cum_twist;
    }
    return NULL;
}

KFO_EXPORT void kfo_session_stats(const kfo_session *s, R *ae, int *tracked, int *iters, int *log_n,
                                  int *counter, int *reset)
{
    *ae = s->last_ae; *tracked = s->last_tracked; *iters = s->iters_run; *log_n = s->log_n;
    *counter = s->counter; *reset = s->reset;
}

/* raycast(): gFusion.cpp:1289-1366 */
KFO_EXPORT void kfo_session_raycast(kfo_session *s)
{
    const kfo_config *c = &s->cfg;
    kfo_raycast(s->vol, c->mode, c->vol, c->dim, s->pose, c->cam, c->nearPlane, c->farPlane,
                step_size(c), KC(0.75) * c->mu, c->width, c->height, s->refV, s->refN);
}

/* Reset(pose, false) + raycast() as the UI does (kinect.cpp:446-477,
 * gFusion.cpp:226-287): zero volume, set pose, arm the 4-frame reset window. */
KFO_EXPORT void kfo_session_reset(kfo_session *s, const R pose[16])
{
    kfo_reset_volume(s->vol, s->cfg.mode, s->cfg.vol);
    s->lib_n = 0; /* poseLibrary.resize(0) */
    /* resetPose (gFusion.cpp:275-286): m_pose_eig = identity rotation with the
     * pose's translation; m_cumTwist = 0 */
    memcpy(s->pose, pose, 16 * sizeof(R));
    for (int i = 0; i < 16; ++i) s->pose_eig[i] = (i % 5 == 0) ? KC(1.0) : KC(0.0);
    s->pose_eig[12] = pose[12]; s->pose_eig[13] = pose[13]; s->pose_eig[14] = pose[14];
    for (int k = 0; k < 6; ++k) s->cum_twist[k] = 0.0;
    kfo_session_raycast(s);
    s->counter = 0;
    s->reset = 1;
}

/* updatePoseFinder (gFusion.cpp:1570-1597): add the pose to the library when
 * its translation is at least 0.2 m (glm::l2Norm) from every entry's. */
static void update_pose_finder(kfo_session *s)
{
    for (int i = 0; i < s->lib_n; ++i) {
        R d0 = s->pose[12] - s->lib[i][12], d1 = s->pose[13] - s->lib[i][13], d2 = s->pose[14] - s->lib[i][14];
        if (KSQRT(dot3(d0, d1, d2, d0, d1, d2)) < KC(0.2)) return;
    }
    if (s->lib_n < KFO_LIB_MAX) memcpy(s->lib[s->lib_n++], s->pose, 16 * sizeof(R));
}

/* depthToVertex(float*) + vertexToNormal(): gFusion.cpp:496-547 */
KFO_EXPORT void kfo_session_depth(kfo_session *s, const float *depth_mm)
{
    const kfo_config *c = &s->cfg;
    size_t n = (size_t)c->width * c->height;
    for (size_t i = 0; i < n; ++i) s->depth[0][i] = (R)depth_mm[i];
    kfo_mip_depth(s->depth[0], c->width, c->height, s->depth[1]);
    kfo_mip_depth(s->depth[1], c->width >> 1, c->height >> 1, s->depth[2]);
    for (int l = 0; l < 3; ++l) {
        R cam_l[4];
        R sc = (R)(1 << l);
        for (int k = 0; k < 4; ++k) cam_l[k] = c->cam[k] / sc;
        kfo_depth_to_vertex(s->depth[l], c->width >> l, c->height >> l, cam_l, s->vert[l]);
        kfo_vertex_to_normal(s->vert[l], c->width >> l, c->height >> l, s->norm[l]);
    }
}

/* Track(): gFusion.cpp:648-740 */
KFO_EXPORT int kfo_session_track(kfo_session *s)
{
    const kfo_config *c = &s->cfg;
    R old[16];
    memcpy(old, s->pose, sizeof(old));
    R ae = 0;
    s->iters_run = 0;
    s->log_n = 0;
    for (int level = 2; level >= 0; --level) {
        int w = c->width >> level, h = c->height >> level;
        for (int it = 0; it < c->iterations[level]; ++it) {
            R view[16];
            kfo_track_view(s->pose, c->cam, view);
            kfo_track(s->vert[level], s->norm[level], w, h, s->refV, s->refN, c->width, c->height,
                      s->pose, view, c->dist_threshold, c->normal_threshold, s->result, s->error, s->J);
            kfo_reduce(s->result, s->error, s->J, w, h, s->partials);
            R b[6], C[36];
            kfo_get_reduction(s->partials, c->rule, b, C, &ae);
            double dC[36], db[6], x[6];
            for (int k = 0; k < 36; ++k) dC[k] = (double)C[k];
            for (int k = 0; k < 6; ++k) db[k] = (double)b[k];
            kfo_solve6(dC, db, x);
            kfo_pose_update(x, s->pose);
            s->iters_run++;
            double nrm = 0;
            for (int k = 0; k < 6; ++k) nrm += x[k] * x[k];
            nrm = sqrt(nrm);
            if (s->log_n < 64) {
                double *L = s->log[s->log_n++];
                L[0] = level; L[1] = (double)ae;
                for (int k = 0; k < 6; ++k) L[2 + k] = x[k];
            }
            if (nrm < 1e-5 && nrm != 0) break;
        }
    }
    int tracked;
    if (ae > KC(0.5) || ae == 0) {
        memcpy(s->pose, old, sizeof(old));
        tracked = 0;
    } else {
        tracked = 1;
        update_pose_finder(s);
    }
    s->last_ae = ae;
    s->last_tracked = tracked;
    return tracked;
}

/* TrackSDF(): gFusion.cpp:835-1038.  Starts from m_pose_eig (not m_pose);
 * per level (coarse to fine) and iteration: camToWorld = exp(Twist(result)) *
 * m_pose_eig_prev, trackSDF + reduceSDF + getSDFReduction, scaled and damped
 * LDLT step; a converged iteration ((AE != 0 && AE < 1e-3) || |step| < 1e-3)
 * marks tracked and ends only its level.  On success m_pose = m_pose_eig =
 * exp(result) * m_pose_eig_prev and m_cumTwist += result; otherwise both are
 * restored. */
KFO_EXPORT int kfo_session_track_sdf(kfo_session *s)
{
    const kfo_config *c = &s->cfg;
    R oldPose[16], prev[16];
    memcpy(oldPose, s->pose, sizeof(oldPose));
    memcpy(prev, s->pose_eig, sizeof(prev));
    double result[6] = {0, 0, 0, 0, 0, 0}, result_prev[6] = {0, 0, 0, 0, 0, 0};
    int tracked = 0;
    s->sdf_log_n = 0;
    memset(s->sdf_log, 0, sizeof(s->sdf_log));
    for (int level = 2; level >= 0; --level) {
        int w = c->width >> level, h = c->height >> level;
        for (int it = 0; it < c->iterations[level]; ++it) {
            R E[16], cam[16];
            kfo_twist_exp(result, E);
            kfo_mul4(E, prev, cam);
            kfo_track_sdf(s->vol, c->mode, c->vol, c->dim, s->vert[level], w, h, cam, s->sdf_buf);
            kfo_reduce_sdf(s->sdf_buf, w, h, s->sdf_partials);
            R b[6], C[36], ae;
            kfo_get_reduction(s->sdf_partials, c->rule, b, C, &ae);
            double cn = kfo_sdf_step(b, C, it, result, result_prev);
            if (s->sdf_log_n < 64) {
                double *L = s->sdf_log[s->sdf_log_n++];
                L[0] = level; L[1] = it; L[2] = (double)ae; L[3] = cn;
                for (int k = 0; k < 6; ++k) L[4 + k] = result[k];
            }
            if (((double)ae != 0.0 && (double)ae < 1e-3) || cn < 1e-3) {
                tracked = 1;
                break;
            }
        }
    }
    double sum = 0.0;
    for (int k = 0; k < 6; ++k) sum += result[k];
    if (sum != sum)
        for (int k = 0; k < 6; ++k) result[k] = 0.0;
    if (tracked) {
        R E[16];
        kfo_twist_exp(result, E);
        kfo_mul4(E, prev, s->pose_eig);
        memcpy(s->pose, s->pose_eig, sizeof(s->pose));
        s->pose[3] = s->pose[7] = s->pose[11] = 0;
        s->pose[15] = KC(1.0);
        for (int k = 0; k < 6; ++k) s->cum_twist[k] += result[k];
    } else {
        memcpy(s->pose_eig, prev, sizeof(prev));
        memcpy(s->pose, oldPose, sizeof(oldPose));
    }
    return tracked;
}

/* recoverPose(): gFusion.cpp:1616-1653 -- for each library entry in order,
 * m_pose = entry, then TrackSDF(); stop at the first success. */
KFO_EXPORT int kfo_session_recover_pose(kfo_session *s)
{
    for (int i = 0; i < s->lib_n; ++i) {
        memcpy(s->pose, s->lib[i], sizeof(s->pose));
        if (kfo_session_track_sdf(s)) return 1;
    }
    return 0;
}

KFO_EXPORT void kfo_session_set_mode(kfo_session *s, int track_mode, int recover)
{
    s->track_mode = track_mode;
    s->recover = recover;
}

KFO_EXPORT int kfo_session_lib_size(const kfo_session *s) { return s->lib_n; }

/* integrate(): gFusion.cpp:1234-1287 */
KFO_EXPORT void kfo_session_integrate(kfo_session *s)
{
    const kfo_config *c = &s->cfg;
    kfo_integrate(s->vol, c->mode, c->vol, c->dim, s->depth[0], c->width, c->height, s->pose, c->cam,
                  c->mu, c->maxWeight);
}

/* The frame block of kinect.cpp:654-695 with integration on: depth ->
 * pyramid, Track and / or TrackSDF (track_mode; P2P alone by default,
 * kinect.h:143-144), integrate if tracked or within the reset window (frames
 * 0-3 after a reset), recoverPose if not tracked, raycast.  Returns tracked. */
KFO_EXPORT int kfo_session_frame(kfo_session *s, const float *depth_mm)
{
    kfo_session_depth(s, depth_mm);
    int tracked = 0;
    if (s->track_mode & 1) tracked = kfo_session_track(s);
    if (s->track_mode & 2) tracked = kfo_session_track_sdf(s);
    s->last_tracked = tracked;
    if (tracked || s->reset) {
        kfo_session_integrate(s);
        if (s->counter > 2) s->reset = 0;
    }
    if (!tracked && s->recover) kfo_session_recover_pose(s);
    s->counter++;
    kfo_session_raycast(s);
    return tracked;
}
