// Helpers shared by the 2.5-D stencil kernels (64-wide tiles staged in LDS, z-march with register queues).
#pragma once

#include "common.h"

namespace tti {

constexpr int BX = 64;

// Cooperative load of one plane tile with halos into LDS. `origin` is the global index of tile element (0, 0), i.e.
// of the point (x0 - HX, y0 - HY, z). Elements whose tile column is >= xlim or row >= ylim are set to zero (they
// are outside the region any valid output reads).
template <int BY, int W, int H>
__device__ __forceinline__ void
load_tile(float (*t)[W], const float* __restrict__ src, long origin, long ldx, int xlim, int ylim)
{
    for (int k = threadIdx.y * BX + threadIdx.x; k < W * H; k += BX * BY) {
        const int ly = k / W;
        const int lx = k - ly * W;
        float v = 0.0f;
        if (lx < xlim && ly < ylim)
            v = src[origin + (long)ly * ldx + lx];
        t[ly][lx] = v;
    }
}

} // namespace tti

namespace tti {
constexpr int BYS = 8;  // rows per block of the reference-structure stencil kernels
constexpr int BZ_DEFAULT = 64;

// Validates the ghost requirement of the radius-4 stencils: the interior must sit at least 4 points inside the
// allocated x and y extents (the z extent is the caller's contract, as in the reference launchers).
inline bool
ranges_ok(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy)
{
    return ixbeg >= R && iybeg >= R && izbeg >= R && ixend >= ixbeg && iyend >= iybeg && izend >= izbeg &&
           ixend + R < ldimx && iyend + R < ldimy;
}

inline dim3
stencil_grid(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int bz)
{
    const int nx = ixend - ixbeg + 1, ny = iyend - iybeg + 1, nz = izend - izbeg + 1;
    return dim3((nx + BX - 1) / BX, (ny + BYS - 1) / BYS, (nz + bz - 1) / bz);
}
} // namespace tti
