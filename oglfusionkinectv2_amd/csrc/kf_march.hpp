// Exact fast-forward of the raycast march (host and device: the host build is
// what the CPU tests check against plain stepping).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace kf {

// Approximate reciprocal for step-count estimates: the hardware reciprocal
// on the device (about 1 ulp), the quotient on the host.  Every estimate
// made with it is verified exactly before use, so the march's result does
// not depend on its rounding.  KF_MARCH_RCP_PERTURB (host test builds only)
// moves the host value by up to 4 ulps to exercise the corrections.
#ifdef KF_MARCH_RCP_PERTURB
float march_rcp_perturbed(float d);
#endif
__host__ __device__ inline float march_rcp(float d)
{
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(d);
#elif defined(KF_MARCH_RCP_PERTURB)
    return march_rcp_perturbed(d);
#else
    return 1.0f / d;
#endif
}

// Empty-space fast-forward: the march's t follows t <- fl(t + s) from the
// given t until the first value >= T (at least one step), leaving t and the
// previous value in t_prev exactly as the loop
//     do { t_prev = t; t += s; } while (t < T);
// would.  Inside one binade [2^(e-1), 2^e) of t every addition adds the same
// amount d = fl(t + s) - t (t is a multiple of the binade's ulp u, so the
// rounding of t + s depends only on s mod u), except when s mod u is exactly
// u / 2 (round-half-even then alternates): runs of additions that stay inside
// the binade are replaced by t + k d, which is a float of that binade and
// comes out of one fma exactly.  Each run takes the first k with
// t + k d >= min(T, 2^e): an estimate from the reciprocal of d, moved by at
// most one and then verified with exact fmas (an estimate that does not
// verify falls back to one plain step).  If t + k d is still inside the
// binade it is the first value >= T; otherwise the k - 1 steps before it stay
// inside and the step after them, which crosses into the next binade, is
// taken plainly, as is every step in the tie case.  Returns the number of
// steps.
__host__ __device__ inline int march_to(float &t, float &t_prev, float s, float T)
{
    int n = 0;
    for (;;) {
        const float tn = t + s;
        int e;
        (void)frexpf(t, &e);  // t in [2^(e-1), 2^e)
        const float top = ldexpf(1.0f, e), u = ldexpf(1.0f, e - 24);
        const float d = tn - t;
        if (t > 0.0f && t < T && tn < top && d > 0.0f && fabsf(s - d) != 0.5f * u) {
            const float lim = fminf(T, top);
            // k >= 1: t + 1 d = tn; the estimate of the first k with t + k d >= lim
            float k = fmaxf(1.0f, ceilf((lim - t) * march_rcp(d)));
            if (k > 1.0f && fmaf(k - 1.0f, d, t) >= lim)
                k -= 1.0f;
            else if (fmaf(k, d, t) < lim)
                k += 1.0f;
            const float tk = fmaf(k, d, t), tk1 = fmaf(k - 1.0f, d, t);
            if (tk >= lim && (k == 1.0f || tk1 < lim) && k < 16777216.0f) {
                if (tk < top) {  // inside the binade, so lim == T: the first value >= T
                    t_prev = tk1;
                    t = tk;
                    return n + (int)k;
                }
                // k >= 2 here (t + d = tn < top): steps 1 .. k - 1 stay
                // below top and T; the next one crosses the binade
                t_prev = fmaf(k - 2.0f, d, t);
                t = tk1;
                n += (int)k - 1;
            }
        }
        // one plain step (binade crossing, rounding tie, or an estimate that
        // did not verify)
        t_prev = t;
        t = t + s;
        ++n;
        if (!(t < T)) return n;
    }
}

}  // namespace kf
