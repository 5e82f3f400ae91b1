// Exact fast-forward of the raycast march (host and device: the host build is
// what the CPU tests check against plain stepping).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace kf {

// Empty-space fast-forward: the march's t follows t <- fl(t + s) from the
// given t until the first value >= T (at least one step), leaving t and the
// previous value in t_prev exactly as the loop
//     do { t_prev = t; t += s; } while (t < T);
// would.  Inside one binade [2^(e-1), 2^e) of t every addition adds the same
// amount d = fl(t + s) - t (t is a multiple of the binade's ulp u, so the
// rounding of t + s depends only on s mod u), except when s mod u is exactly
// u / 2 (round-half-even then alternates): runs of additions that stay inside
// the binade are replaced by t + k d, which is a float of that binade and
// comes out of one fma exactly.  Steps that cross into the next binade, and
// the tie case, are taken one at a time.  Returns the number of steps.
__host__ __device__ inline int march_to(float &t, float &t_prev, float s, float T)
{
    int n = 0;
    for (;;) {
        const float tn = t + s;
        int e;
        (void)frexpf(t, &e);  // t in [2^(e-1), 2^e)
        const float top = ldexpf(1.0f, e), u = ldexpf(1.0f, e - 24);
        const float d = tn - t;
        if (t > 0.0f && tn < top && d > 0.0f && fabsf(s - d) != 0.5f * u) {
            const float rd = 1.0f / d;
            // kb: last k with t + k d < top (>= 1 because t + d = tn < top)
            float kb = fmaxf(1.0f, floorf((top - t) * rd));
            int guard = 0;
            while (kb > 1.0f && fmaf(kb, d, t) >= top && guard++ < 4) kb -= 1.0f;
            while (fmaf(kb + 1.0f, d, t) < top && guard++ < 8) kb += 1.0f;
            if (fmaf(kb, d, t) < top) {
                if (fmaf(kb, d, t) < T) {  // the whole run stays below T
                    t_prev = fmaf(kb - 1.0f, d, t);
                    t = fmaf(kb, d, t);
                    n += (int)kb;
                    continue;  // the next step crosses the binade
                }
                // first k in [1, kb] with t + k d >= T
                float k = fminf(kb, fmaxf(1.0f, ceilf((T - t) * rd)));
                guard = 0;
                while (k > 1.0f && fmaf(k - 1.0f, d, t) >= T && guard++ < 4) k -= 1.0f;
                while (k < kb && fmaf(k, d, t) < T && guard++ < 8) k += 1.0f;
                if (fmaf(k, d, t) >= T && (k == 1.0f || fmaf(k - 1.0f, d, t) < T)) {
                    t_prev = fmaf(k - 1.0f, d, t);
                    t = fmaf(k, d, t);
                    return n + (int)k;
                }
            }
        }
        // one plain step (binade crossing, rounding tie, or estimates that
        // did not settle)
        t_prev = t;
        t = tn;
        ++n;
        if (!(t < T)) return n;
    }
}

}  // namespace kf
