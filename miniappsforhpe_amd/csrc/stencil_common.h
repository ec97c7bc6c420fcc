// Helpers shared by the 2.5-D stencil kernels (64-wide tiles staged in LDS, z-march with register queues).
#pragma once

#include <cstdint>
#include <initializer_list>

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

// The quads (4 consecutive x points) of a W x H plane tile that one thread of an NT-thread block moves: quad
// q = tid + k NT sits at tile row q / (W/4), column 4 (q mod W/4). Tile elements at columns >= xlim or rows >= ylim
// read as zero (no valid output reads them); quads wholly outside are not loaded, so no load leaves the rows of the
// allocation. The launchers check that ldx is a multiple of 4 and the volumes are 16-byte aligned, and every tile
// origin is a multiple of 4 columns, so each quad is one aligned float4 load and one aligned LDS store. fetch()
// issues the loads; commit() masks and stores them a plane later (a select right after a load would wait for it).
template <int W, int H, int NT>
struct TileQuads
{
    static constexpr int QPR = W / 4, NQ = QPR * H, QT = (NQ + NT - 1) / NT;
    long off[QT]; // element offset from the tile origin within a plane, -1: not loaded
    int lds[QT];  // first element of the quad in the W x H LDS tile, -1: no quad for this slot
    int keep[QT]; // elements of the quad inside xlim (0..4)

    __device__ TileQuads(int tid, long ldx, int xlim, int ylim)
    {
#pragma unroll
        for (int k = 0; k < QT; ++k) {
            const int q = tid + k * NT;
            const int row = q / QPR, lx = 4 * (q - row * QPR);
            const bool has = q < NQ;
            lds[k] = has ? row * W + lx : -1;
            keep[k] = has && row < ylim ? min(max(xlim - lx, 0), 4) : 0;
            off[k] = keep[k] > 0 ? (long)row * ldx + lx : -1;
        }
    }
    __device__ __forceinline__ void fetch(float4 (&r)[QT], const float* __restrict__ origin) const
    {
#pragma unroll
        for (int k = 0; k < QT; ++k)
            if (off[k] >= 0)
                r[k] = *reinterpret_cast<const float4*>(origin + off[k]);
    }
    __device__ __forceinline__ void commit(float* tile, const float4 (&r)[QT]) const
    {
#pragma unroll
        for (int k = 0; k < QT; ++k) {
            if (lds[k] < 0)
                continue;
            const int n = keep[k];
            const float4 v = make_float4(n > 0 ? r[k].x : 0.0f, n > 1 ? r[k].y : 0.0f, n > 2 ? r[k].z : 0.0f,
                                         n > 3 ? r[k].w : 0.0f);
            *reinterpret_cast<float4*>(tile + lds[k]) = v;
        }
    }
};

} // namespace tti

namespace tti {
constexpr int BYS = 8;         // rows per block of the reference-structure stencil kernels
constexpr int BYL = 16;        // rows per block of the light ones (forward derivatives, adjoint update)
constexpr int BZ_DEFAULT = 64; // z planes per block where no choice below applies

// Tile of a stencil block in a (gridDim.x, gridDim.y, gridDim.z) launch in XCD order. The dispatcher hands
// consecutive blocks (x fastest, then y, then z) to the 8 XCDs in turn; here each XCD instead takes a contiguous range
// of tiles in that order (any grid size: the first N mod 8 XCDs take one more), so blocks that run at the same time on
// one XCD are x and y neighbours and share their halo lines in its L2. (Measured at 512^3: derivatives 1.79 -> 1.38 ms
// with 16-row blocks and 256-plane chunks, adjoint derivatives 3.40 -> 3.09 ms, density gradient 3.54 -> 3.46 ms.)
__device__ __forceinline__ void
stencil_tile(int& bx, int& by, int& bz)
{
    const int gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
    const int l = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int q = n / 8, r = n % 8, k = l % 8;
    const int t = k * q + min(k, r) + l / 8;
    bx = t % gx;
    by = (t / gx) % gy;
    bz = t / (gx * gy);
}

// Validates the ghost requirement of the radius-4 stencils: the interior must sit at least 4 points inside the
// allocated x and y extents (the z extent is the caller's contract, as in the reference launchers).
inline bool
ranges_ok(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy)
{
    return ixbeg >= R && iybeg >= R && izbeg >= R && ixend >= ixbeg && iyend >= iybeg && izend >= izbeg &&
           ixend + R < ldimx && iyend + R < ldimy;
}

// Whether the quad-staged kernels apply: tile origins (ixbeg - 4 + 64 k) and rows (ldimx) a multiple of 4 floats and
// every tile source 16-byte aligned, so each tile quad is one aligned float4 load. Otherwise the launchers use the
// element-wise kernels.
inline bool
quads_fit(int ixbeg, int ldimx, std::initializer_list<const float*> tiles)
{
    if (ixbeg % 4 || ldimx % 4)
        return false;
    for (const float* p : tiles)
        if (reinterpret_cast<uintptr_t>(p) % 16)
            return false;
    return true;
}

inline dim3
stencil_grid(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int bz, int by = BYS)
{
    const int nx = ixend - ixbeg + 1, ny = iyend - iybeg + 1, nz = izend - izbeg + 1;
    return dim3((nx + BX - 1) / BX, (ny + by - 1) / by, (nz + bz - 1) / bz);
}

// z planes per block of the XCD-ordered kernels: the deepest of 256 and 128 planes that still gives 512 blocks (two
// per CU), else 64. Deeper chunks recompute fewer halo planes (8 per chunk); at 512^3 the kernels take 256.
inline int
stencil_bz(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int by = BYS)
{
    for (int bz : { 256, 128 }) {
        const dim3 g = stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz, by);
        if ((long)g.x * g.y * g.z >= 512)
            return bz;
    }
    return BZ_DEFAULT;
}
} // namespace tti
