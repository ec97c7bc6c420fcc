// Staggered-grid elastic TTI ("eTTI") operators of the simulator3 driver, written for gfx950.
//
// Fields: velocities f1..f3, stresses f4..f9 (11, 22, 33, 12, 13, 23), strains d4..d9 (11, 12, 13, 22, 23, 33),
// models M1..M21 (the upper triangle of the 6 x 6 stiffness matrix) and M22 (buoyancy-like factor), mixed
// interpolations f10..f21, zprime (per-plane z metric). Layout as every volume here: FP32, x fastest,
// idx = (iz*ldy + iy)*ldx + ix, 4-point ghosts; the interior is [4, 4+n) on each axis.
//
// Kernel shape: blocks of 64 x 8 threads (wave w owns tile row w); x-y plane tiles with 4-point halos are staged in
// LDS, z stencils come from per-thread register queues while the block marches along z (2.5-D blocking).
//
// Coefficients: the reference keeps two 4-entry tables in constant memory (DF_order8: staggered first derivative,
// INT_S_O8: diagonal interpolation; simulator3/simulator3SetUp.cu:25-26) and never fills them. Here they are a
// process-wide host copy set by etti_init_constant and passed to every launch by value (zeros until set).
#include <cstdint>
#include <cstring>

#include "common.h"
#include "stencil_common.h"
#include "tti_etti.h"
#include "tti_hip.h"

namespace tti {
namespace etti {

struct EttiCoef
{
    float df[4]; // D+ g(i) = sum_k df[k] (g(i+k+1) - g(i-k)); D- g(i) = sum_k df[k] (g(i+k) - g(i-k-1))
    float si[4]; // Mix g = 0.5 sum_n si[n-1] sum_{p,q in {n, 1-n}} g(i_a + p - s_a, i_b + q - s_b)
};

static EttiCoef g_etti_coef = {};

constexpr int EBY = 8;   // tile rows per block
constexpr int EBZ = 64;  // output planes per block for the z-marching kernels

// D+ / D- over a row or column of an LDS tile (stride S) or a register queue (q[4] = centre).
template <int S>
__device__ __forceinline__ float
dplus(const float* p, const EttiCoef& c)
{
    float r = c.df[0] * (p[S] - p[0]);
#pragma unroll
    for (int k = 1; k < 4; ++k)
        r = fmaf(c.df[k], p[(k + 1) * S] - p[-k * S], r);
    return r;
}

template <int S>
__device__ __forceinline__ float
dminus(const float* p, const EttiCoef& c)
{
    float r = c.df[0] * (p[0] - p[-S]);
#pragma unroll
    for (int k = 1; k < 4; ++k)
        r = fmaf(c.df[k], p[k * S] - p[(-k - 1) * S], r);
    return r;
}

// Diagonal interpolation in an x-y tile around p (row stride W) with shifts (SX, SY).
template <int W, int SX, int SY>
__device__ __forceinline__ float
mix_xy(const float* p, const EttiCoef& c)
{
    float r = 0.0f;
#pragma unroll
    for (int n = 1; n <= 4; ++n) {
        const int a = n - SX, b = 1 - n - SX, u = (n - SY) * W, v = (1 - n - SY) * W;
        r = fmaf(c.si[n - 1], (p[u + a] + p[u + b]) + (p[v + a] + p[v + b]), r);
    }
    return 0.5f * r;
}

// Tile geometry shared by the kernels: block origin, activity and load limits of one x-y tile.
struct EttiTile
{
    int x0, y0, ix, iy, xlim_h, ylim_h, xlim_c, ylim_c;
    bool active;
    long col;
    __device__ EttiTile(int nx, int ny, long ldx, int bxi, int byi)
    {
        x0 = R + bxi * BX;
        y0 = R + byi * EBY;
        ix = x0 + threadIdx.x;
        iy = y0 + threadIdx.y;
        const int ixend = R + nx - 1, iyend = R + ny - 1;
        active = ix <= ixend && iy <= iyend;
        xlim_h = min(BX + 2 * R, ixend + R + 1 - (x0 - R));
        ylim_h = min(EBY + 2 * R, iyend + R + 1 - (y0 - R));
        xlim_c = min(BX, ixend + 1 - x0);
        ylim_c = min(EBY, iyend + 1 - y0);
        col = (long)iy * ldx + ix;
    }
};

// ------------------------------------------------------------------------------------------------------------------
// Velocity update (vPlainKernel, simulator3/simulator3SetUp.cu:820-975):
//   f1 += 2 M22(x-1) (i1 D-x f4 + i2 D+y f7 + zp D+z f8)
//   f2 += 2 M22(y-1) (i1 D+x f7 + i2 D-y f5 + zp D+z f9)
//   f3 += 2 M22(z-1) (i1 D+x f8 + i2 D+y f9 + zp D-z f6)
// ------------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BX* EBY) void k_etti_velocity(
    float* __restrict__ f1, float* __restrict__ f2, float* __restrict__ f3, const float* __restrict__ f4,
    const float* __restrict__ f5, const float* __restrict__ f6, const float* __restrict__ f7,
    const float* __restrict__ f8, const float* __restrict__ f9, const float* __restrict__ M22,
    const float* __restrict__ zprime, int nx, int ny, int nz, long ldx, long ldy, int bz, EttiCoef c, float i1,
    float i2)
{
    // Per plane z (LDS buffer b): the tiles of plane z + 1, loaded during plane z - 1, are stored into buffer 1 - b
    // (last read before the previous barrier), the tiles of plane z + 2 and the column words of plane z + 5 are
    // loaded, plane z is computed from buffer b; one barrier per plane.
    constexpr int W = BX + 2 * R, H = EBY + 2 * R, NT = BX * EBY;
    __shared__ __align__(16) float t4[2][EBY][W]; // x halo
    __shared__ __align__(16) float t8[2][EBY][W];
    __shared__ __align__(16) float t5[2][H][BX]; // y halo
    __shared__ __align__(16) float t9[2][H][BX];
    __shared__ __align__(16) float t7[2][H][W]; // x and y halos
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi); // XCD order, as the type2 kernels
    const EttiTile T(nx, ny, ldx, bxi, byi);
    const int z0 = R + bzi * bz, z1 = min(z0 + bz, R + nz);
    const long sz = ldx * ldy;
    const TileQuads<W, EBY, NT> qx(tid, ldx, T.xlim_h, T.ylim_c);
    const TileQuads<BX, H, NT> qy(tid, ldx, T.xlim_c, T.ylim_h);
    const TileQuads<W, H, NT> qxy(tid, ldx, T.xlim_h, T.ylim_h);
    const long ox = (long)T.y0 * ldx + (T.x0 - R), oy = (long)(T.y0 - R) * ldx + T.x0,
               oxy = (long)(T.y0 - R) * ldx + (T.x0 - R);
    float4 r4[qx.QT], r8[qx.QT], r5[qy.QT], r9[qy.QT], r7[qxy.QT];
    auto fetch = [&](int z) {
        const long pz = (long)z * sz;
        qx.fetch(r4, f4 + pz + ox);
        qx.fetch(r8, f8 + pz + ox);
        qy.fetch(r5, f5 + pz + oy);
        qy.fetch(r9, f9 + pz + oy);
        qxy.fetch(r7, f7 + pz + oxy);
    };
    auto commit = [&](int b) {
        qx.commit(&t4[b][0][0], r4);
        qx.commit(&t8[b][0][0], r8);
        qy.commit(&t5[b][0][0], r5);
        qy.commit(&t9[b][0][0], r9);
        qxy.commit(&t7[b][0][0], r7);
    };

    float q6[9], q8[9], q9[9]; // q[k] = f(z - 4 + k)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const long i = (long)(z0 - 4 + k) * sz + T.col;
        q6[k] = T.active ? f6[i] : 0.0f;
        q8[k] = T.active ? f8[i] : 0.0f;
        q9[k] = T.active ? f9[i] : 0.0f;
    }
    float n6 = 0.0f, n8 = 0.0f, n9 = 0.0f; // column words of plane z + 4, loaded one plane ahead
    if (T.active) {
        const long i = (long)(z0 + 4) * sz + T.col;
        n6 = f6[i];
        n8 = f8[i];
        n9 = f9[i];
    }
    float m22z = T.active ? M22[(long)(z0 - 1) * sz + T.col] : 0.0f; // M22(x, y, z - 1)
    fetch(z0);
    commit(0);
    if (z0 + 1 < z1)
        fetch(z0 + 1);
    __syncthreads();
    for (int z = z0; z < z1; ++z) {
        const int b = (z - z0) & 1;
        const long i = (long)z * sz + T.col;
        q6[8] = n6;
        q8[8] = n8;
        q9[8] = n9;
        float mx = 0.0f, my = 0.0f, mc = 0.0f, v1 = 0.0f, v2 = 0.0f, v3 = 0.0f;
        if (T.active) {
            if (z + 1 < z1) {
                n6 = f6[i + 5 * sz];
                n8 = f8[i + 5 * sz];
                n9 = f9[i + 5 * sz];
            }
            mx = M22[i - 1];
            my = M22[i - ldx];
            mc = M22[i];
            v1 = f1[i];
            v2 = f2[i];
            v3 = f3[i];
        }
        if (z + 1 < z1)
            commit(b ^ 1);
        if (z + 2 < z1)
            fetch(z + 2);
        if (T.active) {
            const float zp = zprime[z];
            const float a1 = dminus<1>(&t4[b][ty][tx + R], c) * i1 + dplus<W>(&t7[b][ty + R][tx + R], c) * i2 +
                             dplus<1>(&q8[4], c) * zp;
            const float a2 = dplus<1>(&t7[b][ty + R][tx + R], c) * i1 + dminus<BX>(&t5[b][ty + R][tx], c) * i2 +
                             dplus<1>(&q9[4], c) * zp;
            const float a3 = dplus<1>(&t8[b][ty][tx + R], c) * i1 + dplus<BX>(&t9[b][ty + R][tx], c) * i2 +
                             dminus<1>(&q6[4], c) * zp;
            f1[i] = v1 + (2.0f * mx) * a1;
            f2[i] = v2 + (2.0f * my) * a2;
            f3[i] = v3 + (2.0f * m22z) * a3;
            m22z = mc;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            q6[k] = q6[k + 1];
            q8[k] = q8[k + 1];
            q9[k] = q9[k + 1];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Strain (vgPlainKernel, simulator3/simulator3SetUp.cu:1218-1361):
//   d4 = i1 D+x f1          d5 = i2 D-y f1 + i1 D-x f2     d6 = zp D-z f1 + i1 D-x f3
//   d7 = i2 D+y f2          d8 = zp D-z f2 + i2 D-y f3     d9 = zp D+z f3
// ------------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BX* EBY) void k_etti_strain(
    const float* __restrict__ f1, const float* __restrict__ f2, const float* __restrict__ f3, float* __restrict__ d4,
    float* __restrict__ d5, float* __restrict__ d6, float* __restrict__ d7, float* __restrict__ d8,
    float* __restrict__ d9, const float* __restrict__ zprime, int nx, int ny, int nz, long ldx, long ldy, int bz,
    EttiCoef c, float i1, float i2)
{
    // same pipeline as the velocity kernel
    constexpr int W = BX + 2 * R, H = EBY + 2 * R, NT = BX * EBY;
    __shared__ __align__(16) float t1[2][H][W];
    __shared__ __align__(16) float t2[2][H][W];
    __shared__ __align__(16) float t3[2][H][W];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi); // XCD order, as the type2 kernels
    const EttiTile T(nx, ny, ldx, bxi, byi);
    const int z0 = R + bzi * bz, z1 = min(z0 + bz, R + nz);
    const long sz = ldx * ldy;
    const TileQuads<W, H, NT> qt(tid, ldx, T.xlim_h, T.ylim_h);
    const long org = (long)(T.y0 - R) * ldx + (T.x0 - R);
    float4 r1[qt.QT], r2[qt.QT], r3[qt.QT];
    auto fetch = [&](int z) {
        const long pz = (long)z * sz + org;
        qt.fetch(r1, f1 + pz);
        qt.fetch(r2, f2 + pz);
        qt.fetch(r3, f3 + pz);
    };
    auto commit = [&](int b) {
        qt.commit(&t1[b][0][0], r1);
        qt.commit(&t2[b][0][0], r2);
        qt.commit(&t3[b][0][0], r3);
    };

    float q1[9], q2[9], q3[9];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const long i = (long)(z0 - 4 + k) * sz + T.col;
        q1[k] = T.active ? f1[i] : 0.0f;
        q2[k] = T.active ? f2[i] : 0.0f;
        q3[k] = T.active ? f3[i] : 0.0f;
    }
    float n1 = 0.0f, n2 = 0.0f, n3 = 0.0f;
    if (T.active) {
        const long i = (long)(z0 + 4) * sz + T.col;
        n1 = f1[i];
        n2 = f2[i];
        n3 = f3[i];
    }
    fetch(z0);
    commit(0);
    if (z0 + 1 < z1)
        fetch(z0 + 1);
    __syncthreads();
    for (int z = z0; z < z1; ++z) {
        const int b = (z - z0) & 1;
        const long i = (long)z * sz + T.col;
        q1[8] = n1;
        q2[8] = n2;
        q3[8] = n3;
        if (T.active && z + 1 < z1) {
            n1 = f1[i + 5 * sz];
            n2 = f2[i + 5 * sz];
            n3 = f3[i + 5 * sz];
        }
        if (z + 1 < z1)
            commit(b ^ 1);
        if (z + 2 < z1)
            fetch(z + 2);
        if (T.active) {
            const float zp = zprime[z];
            const float* c1 = &t1[b][ty + R][tx + R];
            const float* c2 = &t2[b][ty + R][tx + R];
            const float* c3 = &t3[b][ty + R][tx + R];
            d4[i] = i1 * dplus<1>(c1, c);
            d5[i] = i2 * dminus<W>(c1, c) + i1 * dminus<1>(c2, c);
            d6[i] = zp * dminus<1>(&q1[4], c) + i1 * dminus<1>(c3, c);
            d7[i] = i2 * dplus<W>(c2, c);
            d8[i] = zp * dminus<1>(&q2[4], c) + i2 * dminus<W>(c3, c);
            d9[i] = zp * dplus<1>(&q3[4], c);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            q1[k] = q1[k + 1];
            q2[k] = q2[k + 1];
            q3[k] = q3[k + 1];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Mixed interpolations xz = Mix_xz(g; XS, ZS), yz = Mix_yz(g; YS, ZS) (precompute_xz_yz<XS,YS,ZS>,
// simulator3/simulator3SetUp.cu:47-176). Each plane p is staged once and scattered into the eight pending outputs
// it touches: acc[k] belongs to output plane p + ZS - 4 + k, which reads p at z offset q = 4 - k, weight
// si[n - 1] with n = q for q >= 1 and n = 1 - q otherwise; acc[0] is complete after plane p.
// ------------------------------------------------------------------------------------------------------------------
template <int XS, int YS, int ZS>
__global__ __launch_bounds__(BX* EBY) void k_etti_mixed(const float* __restrict__ g, float* __restrict__ xz,
                                                       float* __restrict__ yz, int nx, int ny, int nz, long ldx,
                                                       long ldy, int bz, EttiCoef c)
{
    // planes p0..p1 are staged in turn (double-buffered, the next plane loaded into registers one plane ahead)
    constexpr int W = BX + 2 * R, H = EBY + 2 * R, NT = BX * EBY;
    __shared__ __align__(16) float t[2][H][W];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi); // XCD order, as the type2 kernels
    const EttiTile T(nx, ny, ldx, bxi, byi);
    const int z0 = R + bzi * bz, z1 = min(z0 + bz, R + nz);
    const long sz = ldx * ldy;
    const TileQuads<W, H, NT> qt(tid, ldx, T.xlim_h, T.ylim_h);
    const long org = (long)(T.y0 - R) * ldx + (T.x0 - R);
    float4 r[qt.QT];
    float axz[8], ayz[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        axz[k] = ayz[k] = 0.0f;
    const int p0 = z0 - 3 - ZS, p1 = z1 + 3 - ZS;
    qt.fetch(r, g + (long)p0 * sz + org);
    qt.commit(&t[0][0][0], r);
    qt.fetch(r, g + (long)(p0 + 1) * sz + org);
    __syncthreads();
    for (int p = p0; p <= p1; ++p) {
        const int b = (p - p0) & 1;
        if (p + 1 <= p1)
            qt.commit(&t[b ^ 1][0][0], r);
        if (p + 2 <= p1)
            qt.fetch(r, g + (long)(p + 2) * sz + org);
        const float* row = &t[b][ty + R][tx + R];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int q = 4 - k;
            const int n = q >= 1 ? q : 1 - q;
            const float w = c.si[n - 1];
            axz[k] = fmaf(w, row[n - XS] + row[1 - n - XS], axz[k]);
            ayz[k] = fmaf(w, row[(n - YS) * W] + row[(1 - n - YS) * W], ayz[k]);
        }
        const int zo = p + ZS - 4;
        if (T.active && zo >= z0) {
            const long i = (long)zo * sz + T.col;
            xz[i] = 0.5f * axz[0];
            yz[i] = 0.5f * ayz[0];
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            axz[k] = axz[k + 1];
            ayz[k] = ayz[k + 1];
        }
        axz[7] = ayz[7] = 0.0f;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Stress update (sPlainKernel, simulator3/simulator3SetUp.cu:977-1104): one x-y plane per block z-index; the six
// in-plane mixes come from LDS tiles of the strains, the twelve x-z / y-z mixes from f10..f21.
// ------------------------------------------------------------------------------------------------------------------
struct StressArgs
{
    float* f[6];        // f4..f9
    const float* M[21]; // M1..M21
    const float* d[6];  // d4..d9
    const float* x[12]; // f10..f21
};

__global__ __launch_bounds__(BX* EBY) void k_etti_stress(StressArgs A, int nx, int ny, long ldx, long ldy,
                                                        EttiCoef c)
{
    constexpr int W = BX + 2 * R, H = EBY + 2 * R, NT = BX * EBY;
    __shared__ __align__(16) float td[6][H][W];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    // dispatch order (one plane per block): the XCD order of the z-marching kernels ran 3% slower here
    const EttiTile T(nx, ny, ldx, blockIdx.x, blockIdx.y);
    const int z = R + blockIdx.z;
    const long sz = ldx * ldy;
    const long i = (long)z * sz + T.col;
    // the pointwise words (21 models, 12 mixes, 6 stresses) are loaded first, so they are in flight while the six
    // strain tiles are staged
    float m[22], x[22], s[6];
    if (T.active) {
#pragma unroll
        for (int k = 1; k <= 21; ++k)
            m[k] = A.M[k - 1][i];
#pragma unroll
        for (int k = 10; k <= 21; ++k)
            x[k] = A.x[k - 10][i];
#pragma unroll
        for (int k = 0; k < 6; ++k)
            s[k] = A.f[k][i];
    }
    const TileQuads<W, H, NT> qt(tid, ldx, T.xlim_h, T.ylim_h);
    const long org = (long)z * sz + (long)(T.y0 - R) * ldx + (T.x0 - R);
    float4 r[6][qt.QT];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        qt.fetch(r[k], A.d[k] + org);
#pragma unroll
    for (int k = 0; k < 6; ++k)
        qt.commit(&td[k][0][0], r[k]);
    __syncthreads();
    if (!T.active)
        return;
    const float* p4 = &td[0][ty + R][tx + R];
    const float* p5 = &td[1][ty + R][tx + R];
    const float* p6 = &td[2][ty + R][tx + R];
    const float* p7 = &td[3][ty + R][tx + R];
    const float* p8 = &td[4][ty + R][tx + R];
    const float* p9 = &td[5][ty + R][tx + R];
    const float d4 = *p4, d5 = *p5, d6 = *p6, d7 = *p7, d8 = *p8, d9 = *p9;
    const float m12 = mix_xy<W, 0, 0>(p5, c); // dv1d2_xx
    const float m11 = mix_xy<W, 1, 1>(p4, c); // dv1d1_xy
    const float m22 = mix_xy<W, 1, 1>(p7, c); // dv2d2_xy
    const float m33 = mix_xy<W, 1, 1>(p9, c); // dv3d3_xy
    const float m23 = mix_xy<W, 1, 0>(p8, c); // dv2d3_xz
    const float m13 = mix_xy<W, 0, 1>(p6, c); // dv1d3_yz
    const float f14 = x[14], f19 = x[19];
    A.f[0][i] = s[0] + (m[1] * d4 + m[2] * d7 + m[3] * d9 + m[6] * m12 + m[5] * f14 + m[4] * f19);
    A.f[1][i] = s[1] + (m[2] * d4 + m[7] * d7 + m[8] * d9 + m[11] * m12 + m[10] * f14 + m[9] * f19);
    A.f[2][i] = s[2] + (m[3] * d4 + m[8] * d7 + m[12] * d9 + m[15] * m12 + m[14] * f14 + m[13] * f19);
    A.f[3][i] = s[3] + (m[6] * m11 + m[11] * m22 + m[15] * m33 + m[20] * x[15] + m[18] * x[18] + m[21] * d5);
    A.f[4][i] = s[4] + (m[5] * x[10] + m[10] * x[16] + m[14] * x[20] + m[20] * x[13] + m[17] * m23 + m[19] * d6);
    A.f[5][i] = s[5] + (m[4] * x[11] + m[9] * x[17] + m[13] * x[21] + m[18] * x[12] + m[17] * m13 + m[16] * d8);
}

// ------------------------------------------------------------------------------------------------------------------
// Pointwise kernels
// ------------------------------------------------------------------------------------------------------------------

// Free surface at izbeg over whole x-y planes: zero at izbeg, odd mirror above (mirror_s_kernel / mirror_v_kernel,
// simulator3/simulator3SetUp.cu:1379-1480). Planes above izbeg read only planes below it.
struct Ptr6
{
    float* p[6];
};

__global__ __launch_bounds__(256) void k_etti_mirror(Ptr6 F, int nf, long plane, int izbeg)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (p >= plane)
        return;
    const long i = (long)iz * plane + p;
    const long m = (long)(2 * izbeg - iz) * plane + p;
    for (int k = 0; k < nf; ++k)
        F.p[k][i] = iz == izbeg ? 0.0f : -F.p[k][m];
}

// g1 (no ghosts) += a1 snap1 + a3 snap3 (update_gv, simulator3/simulator3SetUp.cu:1482-1500)
__global__ __launch_bounds__(256) void k_etti_update_gv(const float* __restrict__ s1, const float* __restrict__ s3,
                                                        float* __restrict__ g, const float* __restrict__ a1,
                                                        const float* __restrict__ a3, int ixbeg, int ixend,
                                                        int iybeg, int izbeg, long sldx, long sldy, long ldx,
                                                        long ldy)
{
    const int ix = ixbeg + blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = iybeg + blockIdx.y;
    const int iz = izbeg + blockIdx.z;
    if (ix > ixend)
        return;
    const long s = ((long)(iz - izbeg) * sldy + (iy - iybeg)) * sldx + (ix - ixbeg);
    const long i = ((long)iz * ldy + iy) * ldx + ix;
    g[s] += a1[i] * s1[s] + a3[i] * s3[s];
}

// p = -(2 f4 + 3 f5 + 4 f6 + 5 f7 + 6 f8 + 7 f9) / 27 (compute_p_kernel, simulator3/simulator3SetUp.cu:273-287)
__global__ __launch_bounds__(256) void k_etti_p(long n, float* __restrict__ p, const float* __restrict__ f4,
                                                const float* __restrict__ f5, const float* __restrict__ f6,
                                                const float* __restrict__ f7, const float* __restrict__ f8,
                                                const float* __restrict__ f9)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    p[i] = -(2.0f * f4[i] + 3.0f * f5[i] + 4.0f * f6[i] + 5.0f * f7[i] + 6.0f * f8[i] + 7.0f * f9[i]) / 27.0f;
}

// Model fill of init_cijs_kernel (simulator3/simulator3SetUp.cu:230-271): the first nvol of the 28 volumes (22 or 28)
// get cos/sin of weighted index sums over x in [ixbeg, ixend], every y and z; zprime[z] = cos(6 z).
struct Ptr28
{
    float* p[28];
};

__constant__ const signed char k_cijs_w[28][4] = {
    // {cos = 0 / sin = 1, w_z, w_y, w_x}: M1..M22, then d4..d9
    { 0, 2, 1, 1 }, { 1, 2, 1, 1 }, { 0, 1, 2, 1 }, { 1, 1, 2, 1 }, { 0, 1, 1, 2 }, { 1, 1, 1, 2 },
    { 0, 3, 1, 1 }, { 1, 3, 1, 1 }, { 0, 1, 3, 1 }, { 1, 1, 3, 1 }, { 0, 1, 1, 3 }, { 1, 1, 1, 3 },
    { 0, 4, 1, 1 }, { 1, 4, 1, 1 }, { 0, 1, 4, 1 }, { 1, 1, 4, 1 }, { 0, 1, 1, 4 }, { 1, 1, 1, 4 },
    { 0, 5, 1, 1 }, { 1, 5, 1, 1 }, { 0, 1, 5, 1 }, { 1, 1, 5, 1 }, { 0, 1, 1, 5 }, { 1, 1, 1, 5 },
    { 0, 6, 1, 1 }, { 1, 6, 1, 1 }, { 0, 1, 6, 1 }, { 1, 1, 6, 1 },
};

__global__ __launch_bounds__(256) void k_etti_init_cijs(Ptr28 V, int nvol, float* __restrict__ zprime, int ixbeg,
                                                        int ixend, long ldx, long ldy)
{
    const int ix = ixbeg + blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y, iz = blockIdx.z;
    if (ix == ixbeg && iy == 0)
        zprime[iz] = cosf(6.0f * iz);
    if (ix > ixend)
        return;
    const long i = ((long)iz * ldy + iy) * ldx + ix;
    for (int k = 0; k < nvol; ++k) {
        const float a = (float)k_cijs_w[k][1] * iz + (float)k_cijs_w[k][2] * iy + (float)k_cijs_w[k][3] * ix;
        V.p[k][i] = k_cijs_w[k][0] ? sinf(a) : cosf(a);
    }
}

// Point source of init_src_kernel (simulator3/simulator3SetUp.cu:289-300): f(3 + k) = k * p at one index.
__global__ void k_etti_init_src(Ptr6 F, long i, float p)
{
    for (int k = 0; k < 6; ++k)
        F.p[k][i] = p * (float)(k + 1);
}

} // namespace etti
} // namespace tti

// ======================================================================================================================
// C ABI (declared in include/tti_etti.h)
// ======================================================================================================================
using namespace tti;
using namespace tti::etti;

namespace {

bool
dims_ok(int nx, int ny, int nz, int ldimx, int ldimy)
{
    return nx > 0 && ny > 0 && nz > 0 && ldimx >= nx + 2 * R && ldimy >= ny + 2 * R;
}

bool
all_set(const float* const* p, int n)
{
    for (int k = 0; k < n; ++k)
        if (!p[k])
            return false;
    return true;
}

// The stencil kernels stage tiles with aligned float4 loads: rows a multiple of 4 floats, volumes 16-byte aligned.
bool
quads_ok(int ldimx, const float* const* p, int n)
{
    if (ldimx % 4)
        return false;
    for (int k = 0; k < n; ++k)
        if (reinterpret_cast<uintptr_t>(p[k]) % 16)
            return false;
    return true;
}

dim3
etti_grid(int nx, int ny, int nz, int bz)
{
    return dim3((nx + BX - 1) / BX, (ny + EBY - 1) / EBY, (nz + bz - 1) / bz);
}

// z planes per block of the z-marching kernels: the deepest of 256 and 128 that still gives 512 blocks (two per CU),
// else EBZ (fewer recomputed halo planes per output plane, as the type2 launchers choose). With the XCD order,
// measured at 512^3: velocity 1.70 -> 1.64 ms, strains 1.79 -> 1.76 ms, mixed interpolations 0.54 -> 0.48 ms.
int
etti_bz(int nx, int ny, int nz)
{
    for (int bz : { 256, 128 }) {
        const dim3 g = etti_grid(nx, ny, nz, bz);
        if ((long)g.x * g.y * g.z >= 512)
            return bz;
    }
    return EBZ;
}

// Runs launch() on `stream`; with ms != NULL it also waits for it and adds its event-measured time to *ms, as the
// reference launchers accumulate into their float& arguments.
template <class F>
int
timed_launch(float* ms, void* stream, F&& launch)
{
    const hipStream_t s = (hipStream_t)stream;
    hipEvent_t a = nullptr, b = nullptr;
    if (ms) {
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess || hipEventRecord(a, s) != hipSuccess)
            return launch_status();
    }
    launch();
    int rc = launch_status();
    if (ms && rc == TTI_OK) {
        float t = 0.0f;
        rc = hip_status(hipEventRecord(b, s));
        if (rc == TTI_OK)
            rc = hip_status(hipEventSynchronize(b));
        if (rc == TTI_OK)
            rc = hip_status(hipEventElapsedTime(&t, a, b));
        if (rc == TTI_OK)
            *ms += t;
    }
    if (a)
        (void)hipEventDestroy(a);
    if (b)
        (void)hipEventDestroy(b);
    return rc;
}

template <int XS, int YS, int ZS>
void
launch_mixed(const float* g, float* xz, float* yz, int nx, int ny, int nz, long ldx, long ldy, hipStream_t s)
{
    const int bz = etti_bz(nx, ny, nz);
    k_etti_mixed<XS, YS, ZS><<<etti_grid(nx, ny, nz, bz), dim3(BX, EBY), 0, s>>>(g, xz, yz, nx, ny, nz, ldx, ldy, bz,
                                                                                g_etti_coef);
}

int
mixed(int xs, int ys, int zs, const float* g, float* xz, float* yz, int nx, int ny, int nz, int ldimx, int ldimy,
      void* stream)
{
    const float* p[] = { g, xz, yz };
    if (!all_set(p, 3) || !dims_ok(nx, ny, nz, ldimx, ldimy) || (xs | ys | zs) & ~1 || !quads_ok(ldimx, p, 1))
        return TTI_EINVAL;
    const hipStream_t s = (hipStream_t)stream;
    const long lx = ldimx, ly = ldimy;
    switch (xs * 4 + ys * 2 + zs) {
    case 0: launch_mixed<0, 0, 0>(g, xz, yz, nx, ny, nz, lx, ly, s); break;
    case 1: launch_mixed<0, 0, 1>(g, xz, yz, nx, ny, nz, lx, ly, s); break;
    case 2: launch_mixed<0, 1, 0>(g, xz, yz, nx, ny, nz, lx, ly, s); break;
    case 3: launch_mixed<0, 1, 1>(g, xz, yz, nx, ny, nz, lx, ly, s); break;
    case 4: launch_mixed<1, 0, 0>(g, xz, yz, nx, ny, nz, lx, ly, s); break;
    case 5: launch_mixed<1, 0, 1>(g, xz, yz, nx, ny, nz, lx, ly, s); break;
    case 6: launch_mixed<1, 1, 0>(g, xz, yz, nx, ny, nz, lx, ly, s); break;
    default: launch_mixed<1, 1, 1>(g, xz, yz, nx, ny, nz, lx, ly, s); break;
    }
    return launch_status();
}

} // namespace

extern "C" {

int
etti_init_constant(const float* df8, const float* s8)
{
    if (df8)
        std::memcpy(g_etti_coef.df, df8, sizeof g_etti_coef.df);
    if (s8)
        std::memcpy(g_etti_coef.si, s8, sizeof g_etti_coef.si);
    return TTI_OK;
}

int
etti_get_constant(float* df8, float* s8)
{
    if (df8)
        std::memcpy(df8, g_etti_coef.df, sizeof g_etti_coef.df);
    if (s8)
        std::memcpy(s8, g_etti_coef.si, sizeof g_etti_coef.si);
    return TTI_OK;
}

int
vPlainLoop(float* f1, float* f2, float* f3, const float* f4, const float* f5, const float* f6, const float* f7,
           const float* f8, const float* f9, const float* M22, const float* zprime, int nx, int ny, int nz,
           int ldimx, int ldimy, float* milliseconds_v, void* stream)
{
    const float* p[] = { f1, f2, f3, f4, f5, f6, f7, f8, f9, M22, zprime };
    if (!all_set(p, 11) || !dims_ok(nx, ny, nz, ldimx, ldimy) || !quads_ok(ldimx, p + 3, 6))
        return TTI_EINVAL;
    return timed_launch(milliseconds_v, stream, [&] {
        const int bz = etti_bz(nx, ny, nz);
        k_etti_velocity<<<etti_grid(nx, ny, nz, bz), dim3(BX, EBY), 0, (hipStream_t)stream>>>(
            f1, f2, f3, f4, f5, f6, f7, f8, f9, M22, zprime, nx, ny, nz, ldimx, ldimy, bz, g_etti_coef, 1.01f,
            2.01f);
    });
}

int
vgPlainLoop(const float* f1, const float* f2, const float* f3, float* d4, float* d5, float* d6, float* d7, float* d8,
            float* d9, const float* zprime, int nx, int ny, int nz, int ldimx, int ldimy, float* milliseconds_g,
            void* stream)
{
    const float* p[] = { f1, f2, f3, d4, d5, d6, d7, d8, d9, zprime };
    if (!all_set(p, 10) || !dims_ok(nx, ny, nz, ldimx, ldimy) || !quads_ok(ldimx, p, 3))
        return TTI_EINVAL;
    return timed_launch(milliseconds_g, stream, [&] {
        const int bz = etti_bz(nx, ny, nz);
        k_etti_strain<<<etti_grid(nx, ny, nz, bz), dim3(BX, EBY), 0, (hipStream_t)stream>>>(
            f1, f2, f3, d4, d5, d6, d7, d8, d9, zprime, nx, ny, nz, ldimx, ldimy, bz, g_etti_coef, 1.01f, 2.02f);
    });
}

int
precompute_xz_yz(int xs, int ys, int zs, const float* input, float* drvxz, float* drvyz, int nx, int ny, int nz,
                 int ldimx, int ldimy, void* stream)
{
    return mixed(xs, ys, zs, input, drvxz, drvyz, nx, ny, nz, ldimx, ldimy, stream);
}

int
sPlainKernel_launch(float* const* f4_f9, const float* const* M1_M21, const float* const* d4_d9,
                    const float* const* f10_f21, int nx, int ny, int nz, int ldimx, int ldimy, void* stream)
{
    if (!f4_f9 || !M1_M21 || !d4_d9 || !f10_f21 || !dims_ok(nx, ny, nz, ldimx, ldimy))
        return TTI_EINVAL;
    StressArgs A;
    for (int k = 0; k < 6; ++k) {
        A.f[k] = f4_f9[k];
        A.d[k] = d4_d9[k];
    }
    for (int k = 0; k < 21; ++k)
        A.M[k] = M1_M21[k];
    for (int k = 0; k < 12; ++k)
        A.x[k] = f10_f21[k];
    if (!all_set(A.f, 6) || !all_set(A.d, 6) || !all_set(A.M, 21) || !all_set(A.x, 12) || !quads_ok(ldimx, A.d, 6))
        return TTI_EINVAL;
    k_etti_stress<<<dim3((nx + BX - 1) / BX, (ny + EBY - 1) / EBY, nz), dim3(BX, EBY), 0, (hipStream_t)stream>>>(
        A, nx, ny, ldimx, ldimy, g_etti_coef);
    return launch_status();
}

int
sPlainLoop(float* f4, float* f5, float* f6, float* f7, float* f8, float* f9, const float* M1, const float* M2,
           const float* M3, const float* M4, const float* M5, const float* M6, const float* M7, const float* M8,
           const float* M9, const float* M10, const float* M11, const float* M12, const float* M13,
           const float* M14, const float* M15, const float* M16, const float* M17, const float* M18,
           const float* M19, const float* M20, const float* M21, const float* f1_1, const float* f1_2,
           const float* f1_3, const float* f2_2, const float* f2_3, const float* f3_3, float* f10, float* f11,
           float* f12, float* f13, float* f14, float* f15, float* f16, float* f17, float* f18, float* f19,
           float* f20, float* f21, int nx, int ny, int nz, int ldimx, int ldimy, float* milliseconds_s,
           void* stream)
{
    // the six x-z / y-z mixes (simulator3/simulator3SetUp.cu:1186-1191), then the stress update (:1193-1214); only
    // the update is timed, as in the reference. Every pointer is checked before the first launch.
    float* f[6] = { f4, f5, f6, f7, f8, f9 };
    const float* M[21] = { M1, M2, M3, M4, M5, M6, M7, M8, M9, M10, M11, M12, M13, M14, M15, M16, M17, M18, M19,
                           M20, M21 };
    const float* d[6] = { f1_1, f1_2, f1_3, f2_2, f2_3, f3_3 };
    const float* x[12] = { f10, f11, f12, f13, f14, f15, f16, f17, f18, f19, f20, f21 };
    if (!all_set(f, 6) || !all_set(M, 21) || !all_set(d, 6) || !all_set(x, 12) || !dims_ok(nx, ny, nz, ldimx, ldimy))
        return TTI_EINVAL;
    const struct
    {
        const float* in;
        float *xz, *yz;
        int xs, ys, zs;
    } pre[6] = { { f1_1, f10, f11, 1, 1, 1 }, { f2_2, f16, f17, 1, 1, 1 }, { f3_3, f20, f21, 1, 1, 1 },
                 { f1_2, f12, f13, 0, 0, 1 }, { f1_3, f14, f15, 0, 1, 0 }, { f2_3, f18, f19, 1, 0, 0 } };
    for (const auto& q : pre) {
        const int rc = mixed(q.xs, q.ys, q.zs, q.in, q.xz, q.yz, nx, ny, nz, ldimx, ldimy, stream);
        if (rc != TTI_OK)
            return rc;
    }
    return timed_launch(milliseconds_s, stream,
                        [&] { (void)sPlainKernel_launch(f, M, d, x, nx, ny, nz, ldimx, ldimy, stream); });
}

int
mirror_s_inner(float* f4, float* f5, float* f6, float* f7, float* f8, float* f9, long stride, int izbeg,
               void* stream)
{
    Ptr6 F = { { f4, f5, f6, f7, f8, f9 } };
    if (!all_set(F.p, 6) || stride <= 0 || izbeg < 0)
        return TTI_EINVAL;
    k_etti_mirror<<<dim3((unsigned)((stride + 255) / 256), izbeg + 1), 256, 0, (hipStream_t)stream>>>(F, 6, stride,
                                                                                                       izbeg);
    return launch_status();
}

int
mirror_v_inner(float* f1, float* f2, float* f3, long stride, int izbeg, void* stream)
{
    Ptr6 F = { { f1, f2, f3, nullptr, nullptr, nullptr } };
    if (!all_set(F.p, 3) || stride <= 0 || izbeg < 0)
        return TTI_EINVAL;
    k_etti_mirror<<<dim3((unsigned)((stride + 255) / 256), izbeg + 1), 256, 0, (hipStream_t)stream>>>(F, 3, stride,
                                                                                                       izbeg);
    return launch_status();
}

int
update_gv_inner(const float* snap_f22, const float* snap_f23, float* g1, const float* a_f22, const float* a_f23,
                int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int snap_ldimx, int snap_ldimy,
                int ldimx, int ldimy, void* stream)
{
    if (!snap_f22 || !snap_f23 || !g1 || !a_f22 || !a_f23 || ixend < ixbeg || iyend < iybeg || izend < izbeg ||
        ixbeg < 0 || iybeg < 0 || izbeg < 0 || snap_ldimx < ixend - ixbeg + 1 || snap_ldimy < iyend - iybeg + 1 ||
        ldimx <= ixend || ldimy <= iyend)
        return TTI_EINVAL;
    k_etti_update_gv<<<dim3((ixend - ixbeg + 256) / 256, iyend - iybeg + 1, izend - izbeg + 1), 256, 0,
                       (hipStream_t)stream>>>(snap_f22, snap_f23, g1, a_f22, a_f23, ixbeg, ixend, iybeg, izbeg,
                                              snap_ldimx, snap_ldimy, ldimx, ldimy);
    return launch_status();
}

int
launch_update_g_r(float* g, const float* f_f24_dx, const float* f_f25_dx, const float* f_f24_dy,
                  const float* f_f25_dy, const float* f_f24_dz, const float* f_f25_dz, const float* f24,
                  const float* f25, const float* m18, const float* m1, const float* m2, const float* m3,
                  const float* m4, const float* rx, const float* ry, const float* rz, int ixbeg, int ixend, int iybeg,
                  int iyend, int izbeg, int izend, int ldimx, int ldimy, int g_ldimx, int g_ldimy, double dx,
                  double dy, double dz, int order, int simple_kernel, void* stream)
{
    // compute_g_r (simulator3/simulator3SetUp.cu:1841-2004) is the density imaging condition of the type2 operator
    // with M18 in the buoyancy slot; the same gfx950 kernel serves both
    return launch_update_rho_gradient_rho(g, f_f24_dx, f_f25_dx, f_f24_dy, f_f25_dy, f_f24_dz, f_f25_dz, f24, f25, m18,
                                          m1, m2, m3, m4, rx, ry, rz, ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx,
                                          ldimy, g_ldimx, g_ldimy, dx, dy, dz, order, simple_kernel, stream);
}

int
etti_init_cijs(float* const* M1_M22, float* const* d4_d9, float* zprime, int ixbeg, int ixend, int ldimx, int ldimy,
               int ldimz, void* stream)
{
    if (!M1_M22 || !zprime || ixbeg < 0 || ixend < ixbeg || ixend >= ldimx || ldimy <= 0 || ldimz <= 0)
        return TTI_EINVAL;
    Ptr28 V = {};
    const int nvol = d4_d9 ? 28 : 22;
    for (int k = 0; k < 22; ++k)
        V.p[k] = M1_M22[k];
    for (int k = 22; k < nvol; ++k)
        V.p[k] = d4_d9[k - 22];
    if (!all_set(V.p, nvol))
        return TTI_EINVAL;
    k_etti_init_cijs<<<dim3((ixend - ixbeg + 256) / 256, ldimy, ldimz), 256, 0, (hipStream_t)stream>>>(
        V, nvol, zprime, ixbeg, ixend, ldimx, ldimy);
    return launch_status();
}

int
etti_compute_p(float* p, const float* f4, const float* f5, const float* f6, const float* f7, const float* f8,
               const float* f9, long n, void* stream)
{
    if (!p || !f4 || !f5 || !f6 || !f7 || !f8 || !f9 || n <= 0)
        return TTI_EINVAL;
    k_etti_p<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(n, p, f4, f5, f6, f7, f8, f9);
    return launch_status();
}

int
etti_init_src(float* f4, float* f5, float* f6, float* f7, float* f8, float* f9, long index, float p, void* stream)
{
    Ptr6 F = { { f4, f5, f6, f7, f8, f9 } };
    if (!all_set(F.p, 6) || index < 0)
        return TTI_EINVAL;
    k_etti_init_src<<<1, 1, 0, (hipStream_t)stream>>>(F, index, p);
    return launch_status();
}

} // extern "C"
