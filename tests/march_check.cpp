// Host check of kf::march_to against plain float stepping.  Prints the number
// of cases and mismatches; exits 1 on any mismatch.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <random>

#include "kf_march.hpp"

#ifdef KF_MARCH_RCP_PERTURB
// the reciprocal moved by up to 4 ulps either way (the device's is about 1
// ulp): the march must still equal plain stepping bit for bit
static std::mt19937 perturb_rng(777);
float kf::march_rcp_perturbed(float d)
{
    float r = 1.0f / d;
    const int k = (int)(perturb_rng() % 9) - 4;
    for (int i = 0; i < k; ++i) r = std::nextafter(r, INFINITY);
    for (int i = 0; i > k; --i) r = std::nextafter(r, 0.0f);
    return r;
}
#endif

static int plain(float &t, float &t_prev, float s, float T)
{
    int n = 0;
    do {
        t_prev = t;
        t += s;
        ++n;
    } while (t < T);
    return n;
}

static bool same(float a, float b) { return std::memcmp(&a, &b, 4) == 0; }

int main(int argc, char **argv)
{
    const long cases = argc > 1 ? std::atol(argv[1]) : 2000000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long bad = 0, ties = 0, crossings = 0;
    for (long i = 0; i < cases; ++i) {
        // t: 1e-3 .. 8 (log-uniform), s: 1e-5 .. 0.2 (log-uniform), several shapes
        float t = (float)std::exp(std::log(1e-3) + U(rng) * (std::log(8.0) - std::log(1e-3)));
        float s = (float)std::exp(std::log(1e-5) + U(rng) * (std::log(0.2) - std::log(1e-5)));
        const int kind = (int)(rng() % 5);
        if (kind == 1) {
            s = std::ldexp(1.0f, -(int)(rng() % 14) - 2);  // power of two steps
        } else if (kind == 2) {
            // s with s mod ulp(t) == ulp(t) / 2: round-half-even ties
            int e;
            (void)std::frexp(t, &e);
            const float u = std::ldexp(1.0f, e - 24);
            const float q = (float)(1 + rng() % 100000);
            s = q * u + 0.5f * u;
            ++ties;
        } else if (kind == 3) {
            // start just below a power of two: the run crosses a binade
            const int e = (int)(rng() % 6) - 3;
            t = std::nextafter(std::ldexp(1.0f, e), 0.0f);
            for (int k = 0; k < (int)(rng() % 50); ++k) t = std::nextafter(t, 0.0f);
            ++crossings;
        }
        const double steps_max = rng() % 10 == 0 ? 200000.0 : 3000.0;  // some long runs over several binades
        const double span = std::exp(U(rng) * std::log(steps_max)) * (double)s;
        float T = (float)((double)t + span * (0.2 + U(rng)));
        if (rng() % 7 == 0) T = t;  // at least one step
        if (rng() % 11 == 0) T = std::nextafter(t + s, INFINITY);
        if (rng() % 13 == 0) T = t - s;  // a target behind t: one step
        float ta = t, pa = 0.0f, tb = t, pb = 0.0f;
        const int na = plain(ta, pa, s, T);
        const int nb = kf::march_to(tb, pb, s, T);
        if (na != nb || !same(ta, tb) || !same(pa, pb)) {
            if (bad < 10)
                std::printf("mismatch t=%.9g s=%.9g T=%.9g: plain n=%d t=%.9g prev=%.9g; jump n=%d t=%.9g prev=%.9g\n",
                            t, s, T, na, ta, pa, nb, tb, pb);
            ++bad;
        }
    }
    std::printf("cases %ld (ties %ld, binade starts %ld) mismatches %ld\n", cases, ties, crossings, bad);
    return bad ? 1 : 0;
}
