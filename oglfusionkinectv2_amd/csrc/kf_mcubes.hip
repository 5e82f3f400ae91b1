// kf_mcubes.hip -- surface export of the TSDF for gfx950: the MCubes pass of
// mcubes.cpp:206-390 (histoPyramids.cs classifyCubes + constructHPLevel,
// traverseHistoPyramids.cs) and the files of exportMesh / exportPointCloud
// (mcubes.cpp:390-509).
//
// The reference builds a histopyramid of per-cube triangle counts and walks
// it from the top for every triangle index; at each level the eight children
// are taken in the order x, then z, then y (cubeOffsets), so triangle t lands
// in the cube whose Morton key (bits interleaved y, z, x per level) holds the
// t-th triangle in key order.  Here the same order comes from a prefix sum:
//   k_mc_count  one wavefront per 8^3 brick (a run of 512 consecutive Morton
//               keys): classify its cubes, write the brick's triangle count;
//               bricks whose dilated occupancy flag is clear hold no value
//               below zero and are skipped exactly (iso <= 0)
//   k_scan_*    exclusive prefix sum of the brick counts in Morton order
//   k_mc_emit   one wavefront per non-empty brick: classify again, prefix
//               over its cubes in Morton order, write every triangle's three
//               vertices at its global index
// so the vertex buffer is the reference's (pos[target * 3 + vertexNr]).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "kf_device.hpp"
#include "kf_session.hpp"
#include "kf_store.hpp"

namespace kf {
namespace {

constexpr int kMcWaves = 4;  // bricks per block (one wavefront each)

struct McArgs {
    const void *tsdf;
    int N[3];
    float iso;
    int levels;              // Morton levels of the brick grid (PB = 1 << levels)
    const uint8_t *dflags;   // dilated brick flags (NULL: classify every brick)
    int NB[3];
    const uint8_t *tri;      // [256][16] edge indices, 255 after the last triangle
    const uint8_t *ntri;     // [256]
    uint32_t *count;         // [PB^3] Morton order
    const uint32_t *offset;  // [PB^3] exclusive prefix of count
    float4 *verts;           // 3 per triangle
    unsigned long long cap;  // triangles verts holds
    unsigned long long nslots;
};

// corner k of the cube in classification-bit order, and edge k's end corners
constexpr int8_t kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1},
                                  {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
constexpr int8_t kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                 {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// kCorner and kEdge as arithmetic (the emit kernel's indices differ per
// lane, where a table lookup would be a memory load per vertex)
__host__ __device__ constexpr int corner_x(int c) { return ((c & 3) == 1 || (c & 3) == 2) ? 1 : 0; }
__host__ __device__ constexpr int corner_y(int c) { return c >> 2; }
__host__ __device__ constexpr int corner_z(int c) { return (c & 3) >= 2 ? 1 : 0; }
__host__ __device__ constexpr int edge_lo(int e) { return e < 8 ? e : e - 8; }
__host__ __device__ constexpr int edge_hi(int e) { return e < 8 ? ((e & 4) | ((e + 1) & 3)) : e - 4; }
constexpr bool corner_edge_formulas_match()
{
    for (int c = 0; c < 8; ++c)
        if (corner_x(c) != kCorner[c][0] || corner_y(c) != kCorner[c][1] || corner_z(c) != kCorner[c][2])
            return false;
    for (int e = 0; e < 12; ++e)
        if (edge_lo(e) != kEdge[e][0] || edge_hi(e) != kEdge[e][1]) return false;
    return true;
}
static_assert(corner_edge_formulas_match(), "corner / edge formulas differ from the tables");

// g[c] for a per-lane index c in 0..7 (selects, no indexed register access)
__device__ __forceinline__ float pick8(const float (&g)[8], int c)
{
    float r = g[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) r = c == i ? g[i] : r;
    return r;
}

// (x, y, z) of a Morton key with `levels` digits; digit = x + 2 z + 4 y
__device__ __forceinline__ void morton_decode(uint32_t m, int levels, int &x, int &y, int &z)
{
    x = y = z = 0;
    for (int b = 0; b < levels; ++b) {
        const uint32_t d = (m >> (3 * b)) & 7u;
        x |= (int)(d & 1u) << b;
        z |= (int)((d >> 1) & 1u) << b;
        y |= (int)(d >> 2) << b;
    }
}

template <class S>
__device__ __forceinline__ float mc_field(const typename S::V *__restrict__ v, const McArgs &a, int x, int y, int z)
{
    if ((unsigned)x >= (unsigned)a.N[0] || (unsigned)y >= (unsigned)a.N[1] || (unsigned)z >= (unsigned)a.N[2])
        return 0.0f;
    return S::raw(v[((size_t)z * a.N[1] + y) * a.N[0] + x]);
}

// A wavefront lane owns the 2x2x2 cubes (x0, y0, z0) + {0,1}^3 (Morton
// digit 0 of its cube keys); they read the 3x3x3 voxels (x0, y0, z0) +
// {0,1,2}^3, loaded once, all loads issued before any is used.
template <class S>
__device__ __forceinline__ void load_block27(const typename S::V *__restrict__ v, const McArgs &a, int x0, int y0,
                                             int z0, float (&F)[27])
{
#pragma unroll
    for (int i = 0; i < 27; ++i) F[i] = mc_field<S>(v, a, x0 + i % 3, y0 + (i / 3) % 3, z0 + i / 9);
}
__host__ __device__ constexpr int f27(int dx, int dy, int dz) { return dx + 3 * dy + 9 * dz; }

// classifyCubes (histoPyramids.cs:104-170): cube index of the cube at voxel
// (x, y, z) = (x0, y0, z0) + (ox, oy, oz), 0 for the 10-voxel border and for
// cubes without a value > 0 in the first seven corners; its field values
// come from the lane's 3x3x3 block
__device__ __forceinline__ int mc_classify27(const float (&F)[27], const McArgs &a, int x, int y, int z, int ox,
                                             int oy, int oz)
{
    if (x <= 10 || x >= a.N[0] - 10 || y <= 10 || y >= a.N[1] - 10 || z <= 10 || z >= a.N[2] - 10) return 0;
    // field[k] at cubeOffsets[k] = (k & 1, k >> 2, (k >> 1) & 1) in (x, y, z)
    float f[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) f[k] = F[f27(ox + (k & 1), oy + (k >> 2), oz + ((k >> 1) & 1))];
    bool pos = false;
#pragma unroll
    for (int k = 0; k < 7; ++k) pos = pos || f[k] > 0.0f;
    if (!pos) return -1;  // stored as (0, 0): no triangles
    const float iso = a.iso;
    return (int)(f[0] < iso) | (int)(f[1] < iso) << 1 | (int)(f[3] < iso) << 2 | (int)(f[2] < iso) << 3 |
           (int)(f[4] < iso) << 4 | (int)(f[5] < iso) << 5 | (int)(f[7] < iso) << 6 | (int)(f[6] < iso) << 7;
}

__device__ __forceinline__ bool brick_live(const McArgs &a, int bx, int by, int bz)
{
    if (bx * 8 >= a.N[0] || by * 8 >= a.N[1] || bz * 8 >= a.N[2]) return false;
    if (!a.dflags) return true;
    return a.dflags[((size_t)bz * a.NB[1] + by) * a.NB[0] + bx] != 0;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_exclusive_u32(uint32_t v, int lane)
{
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    return inc - v;
}

template <class S>
__global__ __launch_bounds__(64 * kMcWaves) void k_mc_count(McArgs a)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long m = (unsigned long long)blockIdx.x * kMcWaves + (threadIdx.x >> 6);
    if (m >= a.nslots) return;
    int bx, by, bz;
    morton_decode((uint32_t)m, a.levels, bx, by, bz);
    uint32_t n = 0;
    if (brick_live(a, bx, by, bz)) {
        const typename S::V *__restrict__ v = (const typename S::V *)a.tsdf;
        int lx, ly, lz;
        morton_decode((uint32_t)(8 * lane), 3, lx, ly, lz);
        const int x0 = 8 * bx + lx, y0 = 8 * by + ly, z0 = 8 * bz + lz;
        float F[27];
        load_block27<S>(v, a, x0, y0, z0, F);
#pragma unroll
        for (int j = 0; j < 8; ++j) {  // Morton digit j = x + 2 z + 4 y
            const int ox = j & 1, oz = (j >> 1) & 1, oy = j >> 2;
            const int ci = mc_classify27(F, a, x0 + ox, y0 + oy, z0 + oz, ox, oy, oz);
            if (ci > 0) n += a.ntri[ci];
        }
        n = wave_sum_u32(n);
    }
    if (lane == 0) a.count[m] = n;
}

template <class S>
__global__ __launch_bounds__(64 * kMcWaves) void k_mc_emit(McArgs a)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long m = (unsigned long long)blockIdx.x * kMcWaves + (threadIdx.x >> 6);
    if (m >= a.nslots) return;
    if (a.count[m] == 0) return;  // uniform over the wavefront
    int bx, by, bz;
    morton_decode((uint32_t)m, a.levels, bx, by, bz);
    const typename S::V *__restrict__ v = (const typename S::V *)a.tsdf;
    int lx, ly, lz;
    morton_decode((uint32_t)(8 * lane), 3, lx, ly, lz);
    const int x0 = 8 * bx + lx, y0 = 8 * by + ly, z0 = 8 * bz + lz;
    float F[27];
    load_block27<S>(v, a, x0, y0, z0, F);
    int ci[8];
    uint32_t n = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {  // Morton digit j = x + 2 z + 4 y
        const int ox = j & 1, oz = (j >> 1) & 1, oy = j >> 2;
        const int c = mc_classify27(F, a, x0 + ox, y0 + oy, z0 + oz, ox, oy, oz);
        ci[j] = c > 0 ? c : 0;
        n += a.ntri[ci[j]];
    }
    // The wavefront's triangles are offset[m] + 0 .. T - 1, lane l's at
    // excl(l) .. incl(l) - 1, in cube order and then triangle order.  Each
    // pass gives every lane one of them: the owner lane (binary search over
    // the inclusive prefixes), the cube (the owner's per-cube counts, 3 bits
    // each: at most 5 triangles per cube) and the triangle within it; the
    // cube's eight corners are read again (cached: this wavefront has just
    // read them) and the triangle's three vertices written.
    const uint32_t excl = wave_exclusive_u32(n, lane), incl = excl + n;
    const uint32_t T = __shfl(incl, 63, 64);
    const unsigned long long base = a.offset[m];
    uint32_t cnt3 = 0, ci_lo = 0, ci_hi = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        cnt3 |= (uint32_t)a.ntri[ci[j]] << (3 * j);
        if (j < 4) ci_lo |= (uint32_t)ci[j] << (8 * j);
        else ci_hi |= (uint32_t)ci[j] << (8 * (j - 4));
    }
    for (uint32_t p = 0; p < T; p += 64) {  // uniform: every lane takes part in the shuffles
        const uint32_t i = p + (uint32_t)lane;
        int o = 0;  // lanes with incl <= i
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
            const uint32_t v = __shfl(incl, o + step - 1, 64);
            if (v <= i) o += step;
        }
        o = min(o, 63);  // lanes with i >= T (they write nothing)
        const uint32_t oe = __shfl(excl, o, 64), oc = __shfl(cnt3, o, 64);
        const uint32_t olo = __shfl(ci_lo, o, 64), ohi = __shfl(ci_hi, o, 64);
        if (i >= T) continue;
        uint32_t r = i - oe;
        int j = -1;
#pragma unroll
        for (int jj = 0; jj < 8; ++jj) {
            const uint32_t c = (oc >> (3 * jj)) & 7u;
            if (j < 0) {
                if (r < c) j = jj;
                else r -= c;
            }
        }
        const int cube = j < 4 ? (int)((olo >> (8 * j)) & 0xffu) : (int)((ohi >> (8 * (j - 4))) & 0xffu);
        const int k = (int)r;
        const unsigned long long t = base + i;
        if (t >= a.cap) continue;
        int ox0, oy0, oz0;
        morton_decode((uint32_t)(8 * o), 3, ox0, oy0, oz0);
        const int cx = 8 * bx + ox0 + (j & 1), cy = 8 * by + oy0 + (j >> 2), cz = 8 * bz + oz0 + ((j >> 1) & 1);
        float g[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) g[c] = mc_field<S>(v, a, cx + corner_x(c), cy + corner_y(c), cz + corner_z(c));
        const uint4 row = *(const uint4 *)(a.tri + 16 * cube);
        const uint32_t rw[4] = {row.x, row.y, row.z, row.w};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            // traverseHistoPyramids.cs:196-250
            const int b = 3 * k + q;  // byte b of the triangle row
            const uint32_t w = b < 4 ? rw[0] : (b < 8 ? rw[1] : (b < 12 ? rw[2] : rw[3]));
            const int e = (int)((w >> (8 * (b & 3))) & 0xffu);
            const int c0 = edge_lo(e), c1 = edge_hi(e);
            const int p0[3] = {cx + corner_x(c0), cy + corner_y(c0), cz + corner_z(c0)};
            const int p1[3] = {cx + corner_x(c1), cy + corner_y(c1), cz + corner_z(c1)};
            const float v0 = pick8(g, c0);
            float diff = (a.iso - v0) / (pick8(g, c1) - v0);
            if (diff > 1.0f || diff < 0.0f) diff = 0.5f;
            float4 out;
            out.x = (float)p0[0] * (1.0f - diff) + (float)p1[0] * diff;
            out.y = (float)p0[1] * (1.0f - diff) + (float)p1[1] * diff;
            out.z = (float)p0[2] * (1.0f - diff) + (float)p1[2] * diff;
            out.w = 0.0f;
            a.verts[3 * t + q] = out;
        }
    }
}

// Exclusive scan of n uint32 in place of `out` (from `in`): blocks of 1024,
// block totals scanned by one block, then added back.  Grand total to *total.
constexpr int kScanBlock = 1024;

__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t *in, uint32_t *out, uint32_t *part,
                                                     unsigned long long n)
{
    __shared__ uint32_t wsum[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long base = (unsigned long long)blockIdx.x * kScanBlock + 4 * threadIdx.x;
    uint32_t v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    const uint32_t ex = wave_exclusive_u32(s, lane);
    if (lane == 63) wsum[wave] = ex + s;
    __syncthreads();
    uint32_t wbase = 0;
    for (int w = 0; w < wave; ++w) wbase += wsum[w];
    uint32_t run = wbase + ex;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) part[blockIdx.x] = run;
}

// one block: exclusive scan of the nb block totals (nb <= 256 * 16)
__global__ __launch_bounds__(256) void k_scan_top(uint32_t *part, int nb, unsigned long long *total)
{
    __shared__ uint32_t wsum[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kPer = 16;
    uint32_t v[kPer], s = 0;
    const int base = kPer * threadIdx.x;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        v[k] = base + k < nb ? part[base + k] : 0u;
        s += v[k];
    }
    const uint32_t ex = wave_exclusive_u32(s, lane);
    if (lane == 63) wsum[wave] = ex + s;
    __syncthreads();
    uint32_t wbase = 0;
    for (int w = 0; w < wave; ++w) wbase += wsum[w];
    uint32_t run = wbase + ex;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (base + k < nb) part[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) *total = run;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t *out, const uint32_t *part, unsigned long long n)
{
    const unsigned long long base = (unsigned long long)blockIdx.x * kScanBlock + 4 * threadIdx.x;
    const uint32_t add = part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < n) out[base + k] += add;
}

// ---------------------------------------------------------------------------
// Triangle table from the cube topology: the surface crosses every edge whose
// corners lie on different sides of the iso level (bit set = below it); on
// each face the crossings join in pairs, a face with four crossings
// separating its two below-iso corners; the pieces close into polygons, each
// oriented with its normal towards the above-iso side and split into a fan of
// triangles from its lowest edge.  Per-case triangle counts are those of the
// reference's table (nrOfTriangles, tables.h).
// ---------------------------------------------------------------------------
const int kHostCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}};
const int kHostEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                              {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
const int kHostFace[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};

int host_edge(int a, int b)
{
    for (int k = 0; k < 12; ++k)
        if ((kHostEdge[k][0] == a && kHostEdge[k][1] == b) || (kHostEdge[k][0] == b && kHostEdge[k][1] == a)) return k;
    return -1;
}

void build_tables(uint8_t tri[256][16], uint8_t ntri[256])
{
    for (int c = 0; c < 256; ++c) {
        bool below[8];
        for (int v = 0; v < 8; ++v) below[v] = (c >> v) & 1;
        int link[12][2], deg[12] = {0};
        for (const auto &f : kHostFace) {
            int es[4];
            bool cr[4];
            int n = 0;
            for (int i = 0; i < 4; ++i) {
                es[i] = host_edge(f[i], f[(i + 1) % 4]);
                cr[i] = below[f[i]] != below[f[(i + 1) % 4]];
                n += cr[i];
            }
            auto join = [&](int p, int q) {
                link[p][deg[p]++] = q;
                link[q][deg[q]++] = p;
            };
            if (n == 2) {
                int k[2], m = 0;
                for (int i = 0; i < 4; ++i)
                    if (cr[i]) k[m++] = es[i];
                join(k[0], k[1]);
            } else if (n == 4) {
                if (below[f[0]]) {  // corners 0 and 2 are below: cut each off
                    join(es[3], es[0]);
                    join(es[1], es[2]);
                } else {
                    join(es[0], es[1]);
                    join(es[2], es[3]);
                }
            }
        }
        for (int i = 0; i < 16; ++i) tri[c][i] = 255;
        bool used[12] = {false};
        int nt = 0;
        for (int e0 = 0; e0 < 12; ++e0) {
            if (!deg[e0] || used[e0]) continue;
            int cyc[12], m = 0, prev = -1, cur = e0;
            for (;;) {
                cyc[m++] = cur;
                used[cur] = true;
                const int nx = link[cur][0] != prev ? link[cur][0] : link[cur][1];
                if (nx == e0 || used[nx]) break;
                prev = cur;
                cur = nx;
            }
            double P[12][3], nrm[3] = {0, 0, 0};
            for (int i = 0; i < m; ++i)
                for (int k = 0; k < 3; ++k)
                    P[i][k] = 0.5 * (kHostCorner[kHostEdge[cyc[i]][0]][k] + kHostCorner[kHostEdge[cyc[i]][1]][k]);
            for (int i = 0; i < m; ++i) {
                const double *p = P[i], *q = P[(i + 1) % m];
                nrm[0] += (p[1] - q[1]) * (p[2] + q[2]);
                nrm[1] += (p[2] - q[2]) * (p[0] + q[0]);
                nrm[2] += (p[0] - q[0]) * (p[1] + q[1]);
            }
            double dir = 0;  // Newell normal against the below -> above edge directions
            for (int i = 0; i < m; ++i) {
                const int va = kHostEdge[cyc[i]][0], vb = kHostEdge[cyc[i]][1];
                const int lo = below[va] ? va : vb, hi = below[va] ? vb : va;
                for (int k = 0; k < 3; ++k) dir += nrm[k] * (kHostCorner[hi][k] - kHostCorner[lo][k]);
            }
            if (dir < 0)
                for (int i = 1, j = m - 1; i < j; ++i, --j) std::swap(cyc[i], cyc[j]);
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

#define MC_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) return kf_fail(KF_ERR_HIP, (std::string(#call) + ": " + hipGetErrorString(e_)).c_str()); \
    } while (0)

int ensure_mesh_buffers(kf_session *s, unsigned long long slots)
{
    if (!s->mc_tab) {
        static uint8_t tab[256 * 16 + 256];
        static bool built = false;
        if (!built) {
            build_tables(reinterpret_cast<uint8_t(*)[16]>(tab), tab + 256 * 16);
            built = true;
        }
        MC_HIP(hipMalloc((void **)&s->mc_tab, sizeof(tab)));
        MC_HIP(hipMemcpyAsync(s->mc_tab, tab, sizeof(tab), hipMemcpyHostToDevice, s->stream));
        MC_HIP(hipStreamSynchronize(s->stream));
    }
    if (s->mc_slots < slots) {
        (void)hipFree(s->mc_count);
        (void)hipFree(s->mc_offset);
        (void)hipFree(s->mc_part);
        s->mc_count = s->mc_offset = s->mc_part = nullptr;
        s->mc_slots = 0;
        const size_t nb = (slots + kScanBlock - 1) / kScanBlock;
        MC_HIP(hipMalloc((void **)&s->mc_count, slots * 4));
        MC_HIP(hipMalloc((void **)&s->mc_offset, slots * 4));
        MC_HIP(hipMalloc((void **)&s->mc_part, nb * 4));
        s->mc_slots = slots;
    }
    if (!s->mc_total) MC_HIP(hipMalloc((void **)&s->mc_total, sizeof(unsigned long long)));
    return KF_OK;
}

// ---------------------------------------------------------------------------
// Scratch of kf_mesh_generate_noalloc: one file-scope device array per GPU
// (each device the library runs on holds its own copy until the process
// ends), laid out as the triangle table, the scan block totals, and the
// per-brick triangle counts and their prefix sums for up to kScratchSlots
// Morton brick slots (volumes up to 1024 voxels per side).  Host code never
// uses the array's name as an address: set-up takes each device's address
// with hipGetSymbolAddress and keeps it in g_scratch[device].
// ---------------------------------------------------------------------------
constexpr unsigned long long kScratchSlots = 1ull << 21;
constexpr size_t kScratchTab = 0;                                  // 256 x 16 + 256 bytes
constexpr size_t kScratchPart = 4608;                              // block totals, kScratchSlots / 1024 words
constexpr size_t kScratchCount = kScratchPart + (kScratchSlots / kScanBlock) * 4;
constexpr size_t kScratchOffset = kScratchCount + kScratchSlots * 4;
constexpr size_t kScratchBytes = kScratchOffset + kScratchSlots * 4;
static_assert(kScratchCount % 256 == 0 && kScratchOffset % 256 == 0, "scratch regions are 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kScratchMaxDevices] = {};

template <class S>
void launch_mc(const McArgs &a, bool emit, hipStream_t st)
{
    const unsigned blocks = (unsigned)((a.nslots + kMcWaves - 1) / kMcWaves);
    if (emit) hipLaunchKernelGGL(k_mc_emit<S>, dim3(blocks), dim3(64 * kMcWaves), 0, st, a);
    else hipLaunchKernelGGL(k_mc_count<S>, dim3(blocks), dim3(64 * kMcWaves), 0, st, a);
}

void launch_mc_storage(int storage, const McArgs &a, bool emit, hipStream_t st)
{
    switch (storage) {
    case KF_STORAGE_RG16I: launch_mc<StoreRG16I>(a, emit, st); break;
    case KF_STORAGE_F32: launch_mc<StoreF32>(a, emit, st); break;
    default: launch_mc<StoreF16>(a, emit, st); break;
    }
}

// Morton brick slots of a volume's surface export (power-of-two cube of
// 8^3 bricks covering the volume) and its level count
unsigned long long mesh_slots(const int N[3], int &levels)
{
    int P = 8;
    while (P < N[0] || P < N[1] || P < N[2]) P *= 2;
    levels = 0;
    while ((8 << levels) < P) ++levels;
    return 1ull << (3 * levels);
}

const uint8_t *host_tables()
{
    static uint8_t tab[256 * 16 + 256];
    static std::once_flag once;
    std::call_once(once, [] { build_tables(reinterpret_cast<uint8_t(*)[16]>(tab), tab + 256 * 16); });
    return tab;
}

}  // namespace

// Once per device, with that device current: the address of the device's
// scratch array and the triangle table copied into it.  Called by
// kf_session_create, so kf_mesh_generate_noalloc itself neither allocates
// nor synchronises.
int mesh_noalloc_setup(int device)
{
    if (device < 0 || device >= kScratchMaxDevices) return kf_fail(KF_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return KF_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return kf_fail(KF_ERR_HIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
    e = hipMemcpy((unsigned char *)p + kScratchTab, host_tables(), 256 * 16 + 256, hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return kf_fail(KF_ERR_HIP, (std::string("triangle table copy: ") + hipGetErrorString(e)).c_str());
    g_scratch[device] = (unsigned char *)p;
    return KF_OK;
}

}  // namespace kf

using kf::kScanBlock;

extern "C" {

int kf_mc_tables(uint8_t *tri, uint8_t *ntri)
{
    if (!tri || !ntri) return kf_fail(KF_ERR_INVALID, "NULL argument");
    kf::build_tables(reinterpret_cast<uint8_t(*)[16]>(tri), ntri);
    return KF_OK;
}

int kf_mesh_generate(kf_session *s, float iso, uint64_t *triangles)
{
    if (!s) return kf_fail(KF_ERR_INVALID, "session is NULL");
    const int *N = s->cfg.volume_size;
    int levels = 0;
    const unsigned long long slots = kf::mesh_slots(N, levels);
    const size_t nb = (size_t)((slots + kScanBlock - 1) / kScanBlock);
    if (nb > 256 * 16) return kf_fail(KF_ERR_INVALID, "volume too large for the surface export scan");
    int rc = kf::ensure_mesh_buffers(s, slots);
    if (rc) return rc;
    kf::McArgs a{};
    a.tsdf = s->tsdf;
    for (int k = 0; k < 3; ++k) {
        a.N[k] = N[k];
        a.NB[k] = s->NB[k];
    }
    a.iso = iso;
    a.levels = levels;
    // a brick whose dilated flag is clear holds no value below zero around it
    a.dflags = (iso <= 0.0f) ? s->dflags : nullptr;
    a.tri = s->mc_tab;
    a.ntri = s->mc_tab + 256 * 16;
    a.count = s->mc_count;
    a.offset = s->mc_offset;
    a.nslots = slots;
    kf::launch_mc_storage(s->cfg.storage, a, false, s->stream);
    MC_HIP(hipGetLastError());
    hipLaunchKernelGGL(kf::k_scan_blocks, dim3((unsigned)nb), dim3(256), 0, s->stream, s->mc_count, s->mc_offset,
                       s->mc_part, slots);
    hipLaunchKernelGGL(kf::k_scan_top, dim3(1), dim3(256), 0, s->stream, s->mc_part, (int)nb, s->mc_total);
    hipLaunchKernelGGL(kf::k_scan_add, dim3((unsigned)nb), dim3(256), 0, s->stream, s->mc_offset, s->mc_part, slots);
    MC_HIP(hipGetLastError());
    unsigned long long total = 0;
    MC_HIP(hipMemcpyAsync(&total, s->mc_total, sizeof(total), hipMemcpyDeviceToHost, s->stream));
    MC_HIP(hipStreamSynchronize(s->stream));
    if (total > s->mc_cap) {
        (void)hipFree(s->mc_verts);
        s->mc_verts = nullptr;
        s->mc_cap = 0;
        const size_t cap = (size_t)total + total / 4 + 1024;
        MC_HIP(hipMalloc((void **)&s->mc_verts, cap * 3 * sizeof(float4)));
        s->mc_cap = cap;
    }
    a.verts = s->mc_verts;
    a.cap = s->mc_cap;
    if (total) {
        kf::launch_mc_storage(s->cfg.storage, a, true, s->stream);
        MC_HIP(hipGetLastError());
    }
    s->mc_triangles = total;
    if (triangles) *triangles = total;
    return KF_OK;
}

int kf_mesh_generate_noalloc(kf_session *s, float iso, float *vertices_dev, uint64_t max_triangles,
                             uint64_t *triangles_dev)
{
    if (!s || !triangles_dev) return kf_fail(KF_ERR_INVALID, "NULL argument");
    if (max_triangles && !vertices_dev) return kf_fail(KF_ERR_INVALID, "vertices_dev is NULL");
    if ((uintptr_t)vertices_dev % 16 || (uintptr_t)triangles_dev % 8)
        return kf_fail(KF_ERR_INVALID, "vertices_dev must be 16-byte and triangles_dev 8-byte aligned");
    const int *N = s->cfg.volume_size;
    int levels = 0;
    const unsigned long long slots = kf::mesh_slots(N, levels);
    if (slots > kf::kScratchSlots)
        return kf_fail(KF_ERR_INVALID,
                       ("kf_mesh_generate_noalloc: a " + std::to_string(N[0]) + "x" + std::to_string(N[1]) + "x" +
                        std::to_string(N[2]) + " volume needs " + std::to_string(slots) +
                        " brick slots; its scratch holds " + std::to_string(kf::kScratchSlots) +
                        " (volumes up to 1024 voxels per side)")
                           .c_str());
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(kf::g_scratch_mu);
        if (s->device >= 0 && s->device < kf::kScratchMaxDevices) base = kf::g_scratch[s->device];
    }
    if (!base) return kf_fail(KF_ERR_STATE, "surface-export scratch not set up on this device");
    uint32_t *count = reinterpret_cast<uint32_t *>(base + kf::kScratchCount);
    uint32_t *offset = reinterpret_cast<uint32_t *>(base + kf::kScratchOffset);
    uint32_t *part = reinterpret_cast<uint32_t *>(base + kf::kScratchPart);
    const size_t nb = (size_t)((slots + kScanBlock - 1) / kScanBlock);
    kf::McArgs a{};
    a.tsdf = s->tsdf;
    for (int k = 0; k < 3; ++k) {
        a.N[k] = N[k];
        a.NB[k] = s->NB[k];
    }
    a.iso = iso;
    a.levels = levels;
    a.dflags = (iso <= 0.0f) ? s->dflags : nullptr;
    a.tri = base + kf::kScratchTab;
    a.ntri = base + kf::kScratchTab + 256 * 16;
    a.count = count;
    a.offset = offset;
    a.nslots = slots;
    a.verts = reinterpret_cast<float4 *>(vertices_dev);
    a.cap = max_triangles;
    // every slot's count, offset and block total is written before it is read
    kf::launch_mc_storage(s->cfg.storage, a, false, s->stream);
    hipLaunchKernelGGL(kf::k_scan_blocks, dim3((unsigned)nb), dim3(256), 0, s->stream, count, offset, part, slots);
    hipLaunchKernelGGL(kf::k_scan_top, dim3(1), dim3(256), 0, s->stream, part, (int)nb,
                       reinterpret_cast<unsigned long long *>(triangles_dev));
    hipLaunchKernelGGL(kf::k_scan_add, dim3((unsigned)nb), dim3(256), 0, s->stream, offset, part, slots);
    if (max_triangles) kf::launch_mc_storage(s->cfg.storage, a, true, s->stream);
    MC_HIP(hipGetLastError());
    return KF_OK;
}

int kf_mesh_read(kf_session *s, float *vertices, uint64_t max_triangles)
{
    if (!s || !vertices) return kf_fail(KF_ERR_INVALID, "NULL argument");
    const unsigned long long n = s->mc_triangles < max_triangles ? s->mc_triangles : max_triangles;
    if (n) MC_HIP(hipMemcpyAsync(vertices, s->mc_verts, n * 3 * sizeof(float4), hipMemcpyDeviceToHost, s->stream));
    MC_HIP(hipStreamSynchronize(s->stream));
    return KF_OK;
}

// exportMesh (mcubes.cpp:429-509): binary STL, 80-byte header, a triangle
// count of posData.size() / 3 (four times the triangles: posData holds four
// floats per vertex), then per triangle the normal (1, 1, 1), three vertices
// (x, y, z in voxel-index units) and a zero attribute word.
int kf_mesh_export_stl(kf_session *s, const char *path)
{
    if (!s || !path) return kf_fail(KF_ERR_INVALID, "NULL argument");
    const unsigned long long nt = s->mc_triangles;
    std::vector<float> pos((size_t)nt * 12);
    if (nt) {
        int rc = kf_mesh_read(s, pos.data(), nt);
        if (rc) return rc;
    }
    FILE *f = std::fopen(path, "wb");
    if (!f) return kf_fail(KF_ERR_INVALID, (std::string("cannot open ") + path).c_str());
    char hdr[80];
    std::memset(hdr, 0, sizeof(hdr));
    const uint32_t count = (uint32_t)((nt * 12) / 3);
    const uint16_t attr = 0;
    const float one[3] = {1.0f, 1.0f, 1.0f};
    bool ok = std::fwrite(hdr, 1, 80, f) == 80 && std::fwrite(&count, 4, 1, f) == 1;
    std::vector<char> rec(50);
    for (unsigned long long t = 0; t < nt && ok; ++t) {
        const float *p = pos.data() + 12 * t;
        std::memcpy(rec.data(), one, 12);
        std::memcpy(rec.data() + 12, p + 0, 12);
        std::memcpy(rec.data() + 24, p + 4, 12);
        std::memcpy(rec.data() + 36, p + 8, 12);
        std::memcpy(rec.data() + 48, &attr, 2);
        ok = std::fwrite(rec.data(), 1, 50, f) == 50;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return kf_fail(KF_ERR_INVALID, (std::string("write failed: ") + path).c_str());
    return KF_OK;
}

// exportPointCloud (mcubes.cpp:390-427): ASCII PLY of every vertex.
int kf_mesh_export_ply(kf_session *s, const char *path)
{
    if (!s || !path) return kf_fail(KF_ERR_INVALID, "NULL argument");
    const unsigned long long nt = s->mc_triangles;
    std::vector<float> pos((size_t)nt * 12);
    if (nt) {
        int rc = kf_mesh_read(s, pos.data(), nt);
        if (rc) return rc;
    }
    FILE *f = std::fopen(path, "w");
    if (!f) return kf_fail(KF_ERR_INVALID, (std::string("cannot open ") + path).c_str());
    bool ok = std::fprintf(f, "ply\nformat ascii 1.0\nelement vertex %llu\nproperty float x\nproperty float y\n"
                              "property float z\nend_header\n", nt * 3) > 0;
    for (unsigned long long v = 0; v < nt * 3 && ok; ++v)
        ok = std::fprintf(f, "%g %g %g\n", pos[4 * v], pos[4 * v + 1], pos[4 * v + 2]) > 0;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return kf_fail(KF_ERR_INVALID, (std::string("write failed: ") + path).c_str());
    return KF_OK;
}

}  // extern "C"
