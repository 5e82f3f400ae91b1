// kf_synth.cpp -- seeded synthetic depth sequences (stand-in for the Kinect v2
// capture thread of src/interface.cpp, which delivers 512x424 depth in mm).
//
// Scene: inside-out 4 x 3 x 4 m box room, 12 spheres (r in [0.1, 0.4] m) and
// 6 oriented boxes at uniform-random poses, all analytic so depth is the exact
// ray/primitive intersection.  Trajectory: periodic Catmull-Rom spline through
// seeded control points (position + yaw/pitch/roll), re-timed so that every
// frame moves at most 1 cm and 1 degree.  Depth: Kinect v2 pinhole intrinsics,
// Gaussian noise sigma = 1.2 mm + 1.9 mm * (z - 0.4 m)^2, 2 % random dropout,
// invalid (0) beyond 4.5 m, stored as uint16 millimetres.
//
// World frame: the reference's volume is [0, dim]^3 and its initial camera
// pose is translate(dim/2, dim/2, 0) looking down +z with image y pointing to
// +y (kinect.cpp:37).  The room is placed so that the back wall (z = 1.9) and
// the floor (y = 1.9) fall inside a 2 m cube.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

namespace {

struct Sphere {
    double c[3], r;
};
struct Box {
    double c[3], R[3][3], h[3];  // centre, rotation (columns = box axes), half extents
};

struct Scene {
    double room_lo[3], room_hi[3];
    std::vector<Sphere> spheres;
    std::vector<Box> boxes;
};

const double kRoomLo[3] = {-1.0, -1.1, -2.1};
const double kRoomHi[3] = {3.0, 1.9, 1.9};

void rot_from_quat(double w, double x, double y, double z, double R[3][3])
{
    const double n = std::sqrt(w * w + x * x + y * y + z * z);
    w /= n; x /= n; y /= n; z /= n;
    R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - w * z); R[0][2] = 2 * (x * z + w * y);
    R[1][0] = 2 * (x * y + w * z); R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - w * x);
    R[2][0] = 2 * (x * z - w * y); R[2][1] = 2 * (y * z + w * x); R[2][2] = 1 - 2 * (x * x + y * y);
}

void make_scene(uint64_t seed, Scene &s)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> Nd(0.0, 1.0);
    for (int k = 0; k < 3; ++k) {
        s.room_lo[k] = kRoomLo[k];
        s.room_hi[k] = kRoomHi[k];
    }
    auto centre = [&](double c[3]) {
        c[0] = 0.0 + 2.0 * U(rng);
        c[1] = 0.3 + 1.6 * U(rng);
        c[2] = 0.9 + 0.9 * U(rng);
    };
    for (int i = 0; i < 12; ++i) {
        Sphere sp;
        centre(sp.c);
        sp.r = 0.1 + 0.3 * U(rng);
        s.spheres.push_back(sp);
    }
    for (int i = 0; i < 6; ++i) {
        Box b;
        centre(b.c);
        for (int k = 0; k < 3; ++k) b.h[k] = 0.05 + 0.2 * U(rng);
        const double qw = Nd(rng), qx = Nd(rng), qy = Nd(rng), qz = Nd(rng);
        rot_from_quat(qw, qx, qy, qz, b.R);
        s.boxes.push_back(b);
    }
}

// nearest positive hit distance along o + t d, or +inf
double intersect(const Scene &s, const double o[3], const double d[3])
{
    double best = INFINITY;
    // room interior: exit distance
    {
        double texit = INFINITY;
        for (int k = 0; k < 3; ++k) {
            if (d[k] > 0) texit = std::min(texit, (s.room_hi[k] - o[k]) / d[k]);
            else if (d[k] < 0) texit = std::min(texit, (s.room_lo[k] - o[k]) / d[k]);
        }
        if (texit > 0) best = texit;
    }
    for (const Sphere &sp : s.spheres) {
        const double oc[3] = {o[0] - sp.c[0], o[1] - sp.c[1], o[2] - sp.c[2]};
        const double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
        const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - sp.r * sp.r;
        const double disc = b * b - a * c;
        if (disc < 0) continue;
        const double sq = std::sqrt(disc);
        double t = (-b - sq) / a;
        if (t <= 1e-9) t = (-b + sq) / a;
        if (t > 1e-9 && t < best) best = t;
    }
    for (const Box &bx : s.boxes) {
        double lo[3], ld[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = 0;
            ld[k] = 0;
            for (int m = 0; m < 3; ++m) {
                lo[k] += bx.R[m][k] * (o[m] - bx.c[m]);
                ld[k] += bx.R[m][k] * d[m];
            }
        }
        double tn = -INFINITY, tf = INFINITY;
        bool miss = false;
        for (int k = 0; k < 3; ++k) {
            if (std::fabs(ld[k]) < 1e-15) {
                if (lo[k] < -bx.h[k] || lo[k] > bx.h[k]) miss = true;
                continue;
            }
            double t1 = (-bx.h[k] - lo[k]) / ld[k], t2 = (bx.h[k] - lo[k]) / ld[k];
            if (t1 > t2) std::swap(t1, t2);
            tn = std::max(tn, t1);
            tf = std::min(tf, t2);
        }
        if (miss || tn > tf) continue;
        const double t = tn > 1e-9 ? tn : tf;
        if (t > 1e-9 && t < best) best = t;
    }
    return best;
}

// --- trajectory -------------------------------------------------------------
struct Key {
    double p[3], a[3];  // position, (yaw, pitch, roll) radians
};

void rot_ypr(const double a[3], double R[3][3])
{
    // R = Ry(yaw) * Rx(pitch) * Rz(roll)
    const double cy = std::cos(a[0]), sy = std::sin(a[0]);
    const double cp = std::cos(a[1]), sp = std::sin(a[1]);
    const double cr = std::cos(a[2]), sr = std::sin(a[2]);
    const double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
    const double Rx[3][3] = {{1, 0, 0}, {0, cp, -sp}, {0, sp, cp}};
    const double Rz[3][3] = {{cr, -sr, 0}, {sr, cr, 0}, {0, 0, 1}};
    double T[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T[i][j] = 0;
            for (int k = 0; k < 3; ++k) T[i][j] += Ry[i][k] * Rx[k][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0;
            for (int k = 0; k < 3; ++k) R[i][j] += T[i][k] * Rz[k][j];
        }
}

double catmull(double p0, double p1, double p2, double p3, double t)
{
    const double t2 = t * t, t3 = t2 * t;
    return 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t2 + (-p0 + 3 * p1 - 3 * p2 + p3) * t3);
}

struct Spline {
    std::vector<Key> k;
    void eval(double u, double p[3], double a[3]) const
    {
        const int n = (int)k.size();
        double fl = std::floor(u);
        int i = (int)fl;
        const double t = u - fl;
        auto at = [&](int j) -> const Key & { return k[((j % n) + n) % n]; };
        const Key &k0 = at(i - 1), &k1 = at(i), &k2 = at(i + 1), &k3 = at(i + 2);
        for (int m = 0; m < 3; ++m) {
            p[m] = catmull(k0.p[m], k1.p[m], k2.p[m], k3.p[m], t);
            a[m] = catmull(k0.a[m], k1.a[m], k2.a[m], k3.a[m], t);
        }
    }
};

void make_spline(uint64_t seed, Spline &sp)
{
    std::mt19937_64 rng(seed ^ 0x9e3779b97f4a7c15ULL);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double deg = M_PI / 180.0;
    Key k0 = {{1.0, 1.0, 0.0}, {0, 0, 0}};
    sp.k.push_back(k0);
    for (int i = 0; i < 8; ++i) {
        Key k;
        k.p[0] = 1.0 + 0.2 * U(rng);
        k.p[1] = 1.0 + 0.15 * U(rng);
        k.p[2] = 0.0 + 0.2 * U(rng);
        k.a[0] = 8.0 * deg * U(rng);
        k.a[1] = 8.0 * deg * U(rng);
        k.a[2] = 4.0 * deg * U(rng);
        sp.k.push_back(k);
    }
}

void pose_matrix(const double p[3], const double a[3], float out[16])
{
    double R[3][3];
    rot_ypr(a, R);
    for (int c = 0; c < 3; ++c) {
        for (int r = 0; r < 3; ++r) out[c * 4 + r] = (float)R[r][c];
        out[c * 4 + 3] = 0.0f;
    }
    out[12] = (float)p[0];
    out[13] = (float)p[1];
    out[14] = (float)p[2];
    out[15] = 1.0f;
}

double rot_angle_between(const double a0[3], const double a1[3])
{
    double R0[3][3], R1[3][3];
    rot_ypr(a0, R0);
    rot_ypr(a1, R1);
    double tr = 0;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) tr += R0[k][i] * R1[k][i];
    const double c = std::max(-1.0, std::min(1.0, 0.5 * (tr - 1.0)));
    return std::acos(c);
}

inline uint64_t splitmix(uint64_t x)
{
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
inline double u01(uint64_t h) { return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

}  // namespace

extern "C" {

// Camera-to-world poses (column-major 4x4 float) of `frames` frames.
int kfs_trajectory(uint64_t seed, int frames, float *poses)
{
    if (frames < 0 || (!poses && frames > 0)) return -1;
    Spline sp;
    make_spline(seed, sp);
    const double max_t = 0.0095, max_r = 0.95 * M_PI / 180.0;
    double u = 0.0, du = 0.02;
    double p[3], a[3];
    sp.eval(u, p, a);
    for (int f = 0; f < frames; ++f) {
        pose_matrix(p, a, poses + 16 * f);
        // advance: largest du (doubling / halving) keeping the motion bounded
        double p1[3], a1[3];
        du = std::min(du * 1.5, 0.2);
        for (int tries = 0; tries < 60; ++tries) {
            sp.eval(u + du, p1, a1);
            const double dt = std::sqrt((p1[0] - p[0]) * (p1[0] - p[0]) + (p1[1] - p[1]) * (p1[1] - p[1]) +
                                        (p1[2] - p[2]) * (p1[2] - p[2]));
            const double dr = rot_angle_between(a, a1);
            if (dt <= max_t && dr <= max_r) break;
            du *= 0.5;
        }
        u += du;
        std::memcpy(p, p1, sizeof(p));
        std::memcpy(a, a1, sizeof(a));
    }
    return 0;
}

// Renders depth (uint16 mm) for each pose.  noise = 0 gives the exact depth
// rounded to mm (no noise, no dropout).  Frame f uses noise stream
// (seed, f) so any frame can be re-rendered alone.
int kfs_render(uint64_t seed, const float *poses, int frames, int first_frame, const float cam[4], int w, int h,
               int noise, uint16_t *out)
{
    if (!poses || !cam || !out || w <= 0 || h <= 0 || frames < 0) return -1;
    Scene sc;
    make_scene(seed, sc);
    const long total = (long)frames * h;
#pragma omp parallel for schedule(dynamic, 4)
    for (long fr = 0; fr < total; ++fr) {
        const int f = (int)(fr / h), y = (int)(fr % h);
        const float *P = poses + 16 * f;
        const uint64_t fseed = splitmix(seed * 0x100000001b3ULL + (uint64_t)(first_frame + f));
        for (int x = 0; x < w; ++x) {
            const double dc[3] = {((double)x - cam[2]) / cam[0], ((double)y - cam[3]) / cam[1], 1.0};
            double o[3], d[3];
            for (int r = 0; r < 3; ++r) {
                o[r] = P[12 + r];
                d[r] = P[r] * dc[0] + P[4 + r] * dc[1] + P[8 + r] * dc[2];
            }
            const double t = intersect(sc, o, d);  // camera-z depth because dc.z = 1
            double z = t;
            uint16_t mm = 0;
            if (std::isfinite(z)) {
                if (noise) {
                    const uint64_t hsh = splitmix(fseed ^ ((uint64_t)(y * w + x) * 0xd6e8feb86659fd93ULL));
                    const double u1 = u01(hsh), u2 = u01(splitmix(hsh)), u3 = u01(splitmix(hsh ^ 0x5bd1e995ULL));
                    const double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
                    const double sigma = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
                    z += sigma * g;
                    if (u3 < 0.02) z = 0.0;
                }
                if (z > 0.0 && z <= 4.5) mm = (uint16_t)std::min(65535.0, std::floor(z * 1000.0 + 0.5));
            }
            out[((size_t)f * h + y) * w + x] = mm;
        }
    }
    return 0;
}

// Signed distance of the scene at world points (room interior positive), for
// surface-error statistics.
int kfs_scene_sdf(uint64_t seed, const float *pts, long n, float *sdf)
{
    if (!pts || !sdf || n < 0) return -1;
    Scene sc;
    make_scene(seed, sc);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        double best = INFINITY;
        for (int k = 0; k < 3; ++k) best = std::min(best, std::min(p[k] - sc.room_lo[k], sc.room_hi[k] - p[k]));
        for (const Sphere &sp : sc.spheres) {
            const double dx = p[0] - sp.c[0], dy = p[1] - sp.c[1], dz = p[2] - sp.c[2];
            best = std::min(best, std::sqrt(dx * dx + dy * dy + dz * dz) - sp.r);
        }
        for (const Box &bx : sc.boxes) {
            double q[3], outside = 0.0, inside = -INFINITY;
            for (int k = 0; k < 3; ++k) {
                double l = 0;
                for (int m = 0; m < 3; ++m) l += bx.R[m][k] * (p[m] - bx.c[m]);
                q[k] = std::fabs(l) - bx.h[k];
                outside += std::max(q[k], 0.0) * std::max(q[k], 0.0);
                inside = std::max(inside, q[k]);
            }
            best = std::min(best, std::sqrt(outside) + std::min(inside, 0.0));
        }
        sdf[i] = (float)best;
    }
    return 0;
}

}  // extern "C"
