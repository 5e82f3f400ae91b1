// kf_dilate.hpp -- dilated brick flags and 32^3 coarse cells for the raycast's
// empty-space skipping, shared by k_brick_dilate and the frame loop's
// recoverPose launch (which runs the dilation as its prologue).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kf_internal.hpp"

namespace kf {

// One thread per 32^3 coarse cell c (4^3 bricks): the dilated flag of each of
// its bricks b, dflags(b) = OR of flags over b + {0,1}^3 (a trilinear sample
// whose lower corner lies in brick b reads only voxels of those 8 bricks), and
// cflags(c) = OR of those 64 dilated flags.  The 5x5x5 brick flags the cell
// needs are kept as 25 five-bit row masks.
__device__ __forceinline__ void brick_dilate_body(const BrickArgs &a, int i)
{
    // one wavefront per 32^3 coarse cell: lane l holds brick (4 cx + l % 4,
    // 4 cy + (l / 4) % 4, 4 cz + l / 16); its dilated flag is the OR over the
    // bricks b .. b + 1 in each axis (the trilinear corners of a sample in b)
    const int n = a.NC[0] * a.NC[1] * a.NC[2];
    if (i >= n) return;  // the whole wavefront
    const int lane = threadIdx.x & 63;
    const int cx = i % a.NC[0], cy = (i / a.NC[0]) % a.NC[1], cz = i / (a.NC[0] * a.NC[1]);
    const int bx = 4 * cx + (lane & 3), by = 4 * cy + ((lane >> 2) & 3), bz = 4 * cz + (lane >> 4);
    const bool in = bx < a.NB[0] && by < a.NB[1] && bz < a.NB[2];
    uint32_t d = 0;
    if (in) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int X = bx + (k & 1), Y = by + ((k >> 1) & 1), Z = bz + (k >> 2);
            if (X < a.NB[0] && Y < a.NB[1] && Z < a.NB[2])
                d |= a.flags[((size_t)Z * a.NB[1] + Y) * a.NB[0] + X];
        }
        a.dflags[((size_t)bz * a.NB[1] + by) * a.NB[0] + bx] = d ? 1 : 0;
    }
    const bool any = __any(d != 0);
    if (lane == 0) a.cflags[i] = any ? 1 : 0;
}

// The same for the coarse cells i0 + k stride, k = 0..3, by one wavefront
// (the wavefront's cells are uniform): the 32 flag loads of the four cells are
// issued before any is used (indices clamped into the brick grid, values
// beyond it masked), then the four cells are written as brick_dilate_body
// writes them.
__device__ __forceinline__ void brick_dilate_cells4(const BrickArgs &a, int i0, int stride)
{
    const int n = a.NC[0] * a.NC[1] * a.NC[2];
    const int lane = threadIdx.x & 63;
    uint32_t d[4];
    int bi[4];
    bool in[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = min(i0 + k * stride, n - 1);
        const int cx = i % a.NC[0], cy = (i / a.NC[0]) % a.NC[1], cz = i / (a.NC[0] * a.NC[1]);
        const int bx = 4 * cx + (lane & 3), by = 4 * cy + ((lane >> 2) & 3), bz = 4 * cz + (lane >> 4);
        in[k] = bx < a.NB[0] && by < a.NB[1] && bz < a.NB[2];
        bi[k] = (bz * a.NB[1] + by) * a.NB[0] + bx;
        d[k] = 0u;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int X = bx + (j & 1), Y = by + ((j >> 1) & 1), Z = bz + (j >> 2);
            const bool ok = X < a.NB[0] && Y < a.NB[1] && Z < a.NB[2];
            const int Xc = min(X, a.NB[0] - 1), Yc = min(Y, a.NB[1] - 1), Zc = min(Z, a.NB[2] - 1);
            const uint8_t f = a.flags[((size_t)Zc * a.NB[1] + Yc) * a.NB[0] + Xc];
            d[k] |= ok ? f : 0u;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = i0 + k * stride;
        if (i >= n) break;  // uniform
        if (in[k]) a.dflags[bi[k]] = d[k] ? 1 : 0;
        const bool any = __any(in[k] && d[k] != 0);
        if (lane == 0) a.cflags[i] = any ? 1 : 0;
    }
}

}  // namespace kf
