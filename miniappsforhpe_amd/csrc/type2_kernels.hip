// Variable-density TTI operator (the "type2" family), written for gfx950.
//
// Layout: FP32 volumes, x fastest, idx = (iz*ldy + iy)*ldx + ix, 4-point ghost zones on every side
// (utils/volume_index.h:22, infrastructure/axis.cpp:78-100). Launch ranges are inclusive and in array coordinates,
// exactly as the reference launchers take them.
//
// Kernel shape (all stencil kernels here): a block is 64 x BY threads; wave w owns tile row w, so every LDS access
// of the form tile[y + dy][x + dx] is 64 consecutive words and bank-conflict free. The block marches along z over a
// chunk of `bz` output planes, keeping z-stencil data in per-thread register queues (2.5-D blocking). Work that a
// z-stencil needs "ahead" of the output plane is absorbed from plane p = o + 4 and held in queues until output
// plane o is finalised.
#include "common.h"
#include "order_kernels.h"
#include "type2_internal.h"
#include "stencil_common.h"

namespace tti {

// ------------------------------------------------------------------------------------------------------------------
// a2: d_a f = b * D-_a f  (kernels/kernel_type2_fwd_drv.cu:148-362; launcher :368-409)
// ------------------------------------------------------------------------------------------------------------------
template <int BY>
__global__ __launch_bounds__(BX* BY) void k_fwd_drv(const float* __restrict__ f1, const float* __restrict__ f2,
                                                    const float* __restrict__ b, float* __restrict__ d1x,
                                                    float* __restrict__ d1y, float* __restrict__ d1z,
                                                    float* __restrict__ d2x, float* __restrict__ d2y,
                                                    float* __restrict__ d2z, int ixbeg, int ixend, int iybeg,
                                                    int iyend, int izbeg, int izend, float invdx, float invdy,
                                                    float invdz, long ldx, long ldy, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R;
    __shared__ float t1[H][W];
    __shared__ float t2[H][W];
    const int tx = threadIdx.x, ty = threadIdx.y;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi);
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const long col = (long)iy * ldx + ix;

    float q1[8], q2[8]; // q[k] = f(z - 4 + k)
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        q1[k] = active ? f1[(long)(z0 - 4 + k) * sz + col] : 0.0f;
        q2[k] = active ? f2[(long)(z0 - 4 + k) * sz + col] : 0.0f;
    }
    for (int z = z0; z < z1; ++z) {
        if (active) {
            q1[7] = f1[(long)(z + 3) * sz + col];
            q2[7] = f2[(long)(z + 3) * sz + col];
        }
        const long org = (long)z * sz + (long)(y0 - R) * ldx + (x0 - R);
        load_tile<BY, W, H>(t1, f1, org, ldx, xlim, ylim);
        load_tile<BY, W, H>(t2, f2, org, ldx, xlim, ylim);
        __syncthreads();
        if (active) {
            float ax1 = 0, ay1 = 0, az1 = 0, ax2 = 0, ay2 = 0, az2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                ax1 += c * (t1[ty + R][tx + R + n - 1] - t1[ty + R][tx + R - n]);
                ay1 += c * (t1[ty + R + n - 1][tx + R] - t1[ty + R - n][tx + R]);
                az1 += c * (q1[R + n - 1] - q1[R - n]);
                ax2 += c * (t2[ty + R][tx + R + n - 1] - t2[ty + R][tx + R - n]);
                ay2 += c * (t2[ty + R + n - 1][tx + R] - t2[ty + R - n][tx + R]);
                az2 += c * (q2[R + n - 1] - q2[R - n]);
            }
            const long i = (long)z * sz + col;
            const float bb = b[i];
            d1x[i] = ax1 * invdx * bb;
            d1y[i] = ay1 * invdy * bb;
            d1z[i] = az1 * invdz * bb;
            d2x[i] = ax2 * invdx * bb;
            d2y[i] = ay2 * invdy * bb;
            d2z[i] = az2 * invdz * bb;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            q1[k] = q1[k + 1];
            q2[k] = q2[k + 1];
        }
    }
}

// The same operator with the plane pipeline: the sigma tiles of plane z + 1 are loaded as float4 quads during plane
// z - 1 and stored into the other LDS buffer during plane z; the column words (sigma at z + 4, b at z + 1) one plane
// ahead; one barrier per plane. Needs quads_fit (launcher falls back).
template <int BY>
__global__ __launch_bounds__(BX* BY) void k_fwd_drv_q(const float* __restrict__ f1, const float* __restrict__ f2,
                                                      const float* __restrict__ b, float* __restrict__ d1x,
                                                      float* __restrict__ d1y, float* __restrict__ d1z,
                                                      float* __restrict__ d2x, float* __restrict__ d2y,
                                                      float* __restrict__ d2z, int ixbeg, int ixend, int iybeg,
                                                      int iyend, int izbeg, int izend, float invdx, float invdy,
                                                      float invdz, long ldx, long ldy, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R, NT = BX * BY;
    __shared__ __align__(16) float t1[2][H][W];
    __shared__ __align__(16) float t2[2][H][W];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi);
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const long col = (long)iy * ldx + ix;
    const TileQuads<W, H, NT> qt(tid, ldx, xlim, ylim);
    const long org = (long)(y0 - R) * ldx + (x0 - R);
    float4 r1[qt.QT], r2[qt.QT];
    auto fetch = [&](int z) {
        qt.fetch(r1, f1 + (long)z * sz + org);
        qt.fetch(r2, f2 + (long)z * sz + org);
    };

    float q1[8], q2[8]; // q[k] = f(z - 4 + k)
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        q1[k] = active ? f1[(long)(z0 - 4 + k) * sz + col] : 0.0f;
        q2[k] = active ? f2[(long)(z0 - 4 + k) * sz + col] : 0.0f;
    }
    float n1 = 0.0f, n2 = 0.0f, nb = 0.0f; // sigma(z + 3) and b(z) of the next plane
    if (active) {
        n1 = f1[(long)(z0 + 3) * sz + col];
        n2 = f2[(long)(z0 + 3) * sz + col];
        nb = b[(long)z0 * sz + col];
    }
    fetch(z0);
    qt.commit(&t1[0][0][0], r1);
    qt.commit(&t2[0][0][0], r2);
    if (z0 + 1 < z1)
        fetch(z0 + 1);
    __syncthreads();
    for (int z = z0; z < z1; ++z) {
        const int bf = (z - z0) & 1;
        q1[7] = n1;
        q2[7] = n2;
        const float bb = nb;
        if (active && z + 1 < z1) {
            n1 = f1[(long)(z + 4) * sz + col];
            n2 = f2[(long)(z + 4) * sz + col];
            nb = b[(long)(z + 1) * sz + col];
        }
        if (z + 1 < z1) {
            qt.commit(&t1[bf ^ 1][0][0], r1);
            qt.commit(&t2[bf ^ 1][0][0], r2);
        }
        if (z + 2 < z1)
            fetch(z + 2);
        if (active) {
            const float(*T1)[W] = t1[bf];
            const float(*T2)[W] = t2[bf];
            float ax1 = 0, ay1 = 0, az1 = 0, ax2 = 0, ay2 = 0, az2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                ax1 += c * (T1[ty + R][tx + R + n - 1] - T1[ty + R][tx + R - n]);
                ay1 += c * (T1[ty + R + n - 1][tx + R] - T1[ty + R - n][tx + R]);
                az1 += c * (q1[R + n - 1] - q1[R - n]);
                ax2 += c * (T2[ty + R][tx + R + n - 1] - T2[ty + R][tx + R - n]);
                ay2 += c * (T2[ty + R + n - 1][tx + R] - T2[ty + R - n][tx + R]);
                az2 += c * (q2[R + n - 1] - q2[R - n]);
            }
            const long i = (long)z * sz + col;
            d1x[i] = ax1 * invdx * bb;
            d1y[i] = ay1 * invdy * bb;
            d1z[i] = az1 * invdz * bb;
            d2x[i] = ax2 * invdx * bb;
            d2y[i] = ay2 * invdy * bb;
            d2z[i] = az2 * invdz * bb;
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            q1[k] = q1[k + 1];
            q2[k] = q2[k + 1];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// a3: u += dt*rhs3, w += dt*rhs4 (kernels/kernel_type2_fwd.cu:264-540; launcher :578-665)
//   F_aa = rho * D+_a d_a,  F_ab = X_ab f (1/(4 da db) scaling),  V/H split and model combination.
// ------------------------------------------------------------------------------------------------------------------
template <int BY>
__global__ __launch_bounds__(BX* BY) void k_fwd_update(
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ d1x, const float* __restrict__ d2x,
    const float* __restrict__ d1y, const float* __restrict__ d2y, const float* __restrict__ d1z,
    const float* __restrict__ d2z, const float* __restrict__ f1, const float* __restrict__ f2,
    const float* __restrict__ irho, const float* __restrict__ m1, const float* __restrict__ m2,
    const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ rz, int ixbeg, int ixend, int iybeg, int iyend,
    int izbeg, int izend, long ldx, long ldy, float dt, float invdx, float invdy, float invdz, float invdxdy,
    float invdydz, float invdzdx, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R;
    __shared__ float s1[H][W]; // sigma at plane p (halo 4 in x and y)
    __shared__ float s2[H][W];
    __shared__ float x1[BY][W]; // d_x at plane o (halo 4 in x)
    __shared__ float x2[BY][W];
    __shared__ float y1[H][BX]; // d_y at plane o (halo 4 in y)
    __shared__ float y2[H][BX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi);
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const int xlim_c = min(BX, ixend + 1 - x0);
    const int ylim_c = min(BY, iyend + 1 - y0);
    const long col = (long)iy * ldx + ix;

    // mixed partial sums: Q[j] <-> output plane p - 4 + j while plane p is absorbed
    float qzx1[9], qyz1[9], qzx2[9], qyz2[9];
    float xy1[5], xy2[5]; // F_xy delayed: xy[4] = F_xy(p)
    float dz1[8], dz2[8]; // dz[k] <-> d_z(p - 7 + k)
#pragma unroll
    for (int j = 0; j < 9; ++j)
        qzx1[j] = qyz1[j] = qzx2[j] = qyz2[j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 5; ++j)
        xy1[j] = xy2[j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        dz1[j] = dz2[j] = 0.0f;

    for (int p = z0 - R; p < z1 + R; ++p) {
        const int o = p - R;
        const bool emit = o >= z0;
        {
            const long org = (long)p * sz + (long)(y0 - R) * ldx + (x0 - R);
            load_tile<BY, W, H>(s1, f1, org, ldx, xlim, ylim);
            load_tile<BY, W, H>(s2, f2, org, ldx, xlim, ylim);
        }
        if (emit) {
            const long orgx = (long)o * sz + (long)y0 * ldx + (x0 - R);
            load_tile<BY, W, BY>(x1, d1x, orgx, ldx, xlim, ylim_c);
            load_tile<BY, W, BY>(x2, d2x, orgx, ldx, xlim, ylim_c);
            const long orgy = (long)o * sz + (long)(y0 - R) * ldx + x0;
            load_tile<BY, BX, H>(y1, d1y, orgy, ldx, xlim_c, ylim);
            load_tile<BY, BX, H>(y2, d2y, orgy, ldx, xlim_c, ylim);
        }
        if (active) {
            dz1[7] = d1z[(long)p * sz + col];
            dz2[7] = d2z[(long)p * sz + col];
        }
        __syncthreads();
        if (active) {
            // absorb plane p into the x-z / y-z partial sums (plane offset k = 4 - j)
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int k = R - j;
                if (k != 0) {
                    const int ak = k < 0 ? -k : k;
                    const float e = ecen(ak);
                    qzx1[j] += e * (s1[ty + R][tx + R + k] - s1[ty + R][tx + R - k]);
                    qyz1[j] += e * (s1[ty + R + k][tx + R] - s1[ty + R - k][tx + R]);
                    qzx2[j] += e * (s2[ty + R][tx + R + k] - s2[ty + R][tx + R - k]);
                    qyz2[j] += e * (s2[ty + R + k][tx + R] - s2[ty + R - k][tx + R]);
                }
            }
            float a1 = 0, a2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float e = ecen(n);
                a1 += e * ((s1[ty + R + n][tx + R + n] - s1[ty + R - n][tx + R + n]) +
                           (s1[ty + R - n][tx + R - n] - s1[ty + R + n][tx + R - n]));
                a2 += e * ((s2[ty + R + n][tx + R + n] - s2[ty + R - n][tx + R + n]) +
                           (s2[ty + R - n][tx + R - n] - s2[ty + R + n][tx + R - n]));
            }
            xy1[4] = a1 * invdxdy;
            xy2[4] = a2 * invdxdy;

            if (emit) {
                float fxx1 = 0, fxx2 = 0, fyy1 = 0, fyy2 = 0, fzz1 = 0, fzz2 = 0;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    fxx1 += c * (x1[ty][tx + R + n] - x1[ty][tx + R - n + 1]);
                    fxx2 += c * (x2[ty][tx + R + n] - x2[ty][tx + R - n + 1]);
                    fyy1 += c * (y1[ty + R + n][tx] - y1[ty + R - n + 1][tx]);
                    fyy2 += c * (y2[ty + R + n][tx] - y2[ty + R - n + 1][tx]);
                    fzz1 += c * (dz1[3 + n] - dz1[4 - n]);
                    fzz2 += c * (dz2[3 + n] - dz2[4 - n]);
                }
                const long i = (long)o * sz + col;
                const float rho = 1.0f / irho[i];
                fxx1 *= invdx * rho;
                fxx2 *= invdx * rho;
                fyy1 *= invdy * rho;
                fyy2 *= invdy * rho;
                fzz1 *= invdz * rho;
                fzz2 *= invdz * rho;
                const float fzx1 = qzx1[0] * invdzdx, fzx2 = qzx2[0] * invdzdx;
                const float fyz1 = qyz1[0] * invdydz, fyz2 = qyz2[0] * invdydz;
                const float fxy1 = xy1[0], fxy2 = xy2[0];
                const float vx = rx[i], vy = ry[i], vz = rz[i];
                const float rx2 = vx * vx, ry2 = vy * vy, rz2 = vz * vz;
                const float rxy = 2.0f * vx * vy, ryz = 2.0f * vy * vz, rzx = 2.0f * vz * vx;
                const float V1 = rx2 * fxx1 + ry2 * fyy1 + rz2 * fzz1 + rxy * fxy1 + rzx * fzx1 + ryz * fyz1;
                const float H1 = fxx1 + fyy1 + fzz1 - V1;
                const float V2 = rx2 * fxx2 + ry2 * fyy2 + rz2 * fzz2 + rxy * fxy2 + rzx * fzx2 + ryz * fyz2;
                const float H2 = fxx2 + fyy2 + fzz2 - V2;
                const float mm1 = m1[i], mm2 = m2[i], mm3 = m3[i], mm4 = m4[i];
                const float r3 = mm1 * H1 + mm4 * V1 + mm2 * V2;
                const float r4 = mm2 * H1 + mm3 * V2 + mm4 * H2;
                u[i] += r3 * dt;
                w[i] += r4 * dt;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            qzx1[j] = qzx1[j + 1];
            qyz1[j] = qyz1[j + 1];
            qzx2[j] = qzx2[j + 1];
            qyz2[j] = qyz2[j + 1];
        }
        qzx1[8] = qyz1[8] = qzx2[8] = qyz2[8] = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            xy1[j] = xy1[j + 1];
            xy2[j] = xy2[j + 1];
        }
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            dz1[j] = dz1[j + 1];
            dz2[j] = dz2[j + 1];
        }
    }
}

// The same operator with the plane pipeline of the gfx950 stencils: the tiles of plane p + 1 (sigma, and the d_x / d_y
// tiles of output plane p + 1 - 4) are loaded as float4 quads during plane p - 1 and stored into the other LDS buffer
// during plane p, the d_z column word and the 11 pointwise words of the next output plane are loaded one plane ahead,
// and one barrier per plane separates a buffer's last read from its refill. Needs quads_fit (launcher falls back).
template <int BY>
__global__ __launch_bounds__(BX* BY) void k_fwd_update_q(
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ d1x, const float* __restrict__ d2x,
    const float* __restrict__ d1y, const float* __restrict__ d2y, const float* __restrict__ d1z,
    const float* __restrict__ d2z, const float* __restrict__ f1, const float* __restrict__ f2,
    const float* __restrict__ irho, const float* __restrict__ m1, const float* __restrict__ m2,
    const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ rz, int ixbeg, int ixend, int iybeg, int iyend,
    int izbeg, int izend, long ldx, long ldy, float dt, float invdx, float invdy, float invdz, float invdxdy,
    float invdydz, float invdzdx, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R, NT = BX * BY;
    __shared__ __align__(16) float s1[2][H][W]; // sigma at plane p (halo 4 in x and y)
    __shared__ __align__(16) float s2[2][H][W];
    __shared__ __align__(16) float x1[2][BY][W]; // d_x at plane o (halo 4 in x)
    __shared__ __align__(16) float x2[2][BY][W];
    __shared__ __align__(16) float y1[2][H][BX]; // d_y at plane o (halo 4 in y)
    __shared__ __align__(16) float y2[2][H][BX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi);
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const int xlim_c = min(BX, ixend + 1 - x0);
    const int ylim_c = min(BY, iyend + 1 - y0);
    const long col = (long)iy * ldx + ix;
    const int p0 = z0 - R, pend = z1 + R; // tile planes p in [p0, pend), output o = p - R

    const TileQuads<W, H, NT> qs(tid, ldx, xlim, ylim);
    const TileQuads<W, BY, NT> qx(tid, ldx, xlim, ylim_c);
    const TileQuads<BX, H, NT> qy(tid, ldx, xlim_c, ylim);
    const long os = (long)(y0 - R) * ldx + (x0 - R), ox = (long)y0 * ldx + (x0 - R), oy = (long)(y0 - R) * ldx + x0;
    float4 rs1[qs.QT], rs2[qs.QT], rx1[qx.QT], rx2[qx.QT], ry1[qy.QT], ry2[qy.QT];
    auto fetch = [&](int p) {
        const long pz = (long)p * sz;
        qs.fetch(rs1, f1 + pz + os);
        qs.fetch(rs2, f2 + pz + os);
        if (p - R >= z0) {
            const long oz = (long)(p - R) * sz;
            qx.fetch(rx1, d1x + oz + ox);
            qx.fetch(rx2, d2x + oz + ox);
            qy.fetch(ry1, d1y + oz + oy);
            qy.fetch(ry2, d2y + oz + oy);
        }
    };
    auto commit = [&](int p, int b) {
        qs.commit(&s1[b][0][0], rs1);
        qs.commit(&s2[b][0][0], rs2);
        if (p - R >= z0) {
            qx.commit(&x1[b][0][0], rx1);
            qx.commit(&x2[b][0][0], rx2);
            qy.commit(&y1[b][0][0], ry1);
            qy.commit(&y2[b][0][0], ry2);
        }
    };
    // pointwise words of an output plane: 1/rho, m1..m4, rx, ry, rz, u, w
    float pw[10];
    auto load_pw = [&](int o) {
        if (active && o >= z0 && o < z1) {
            const long i = (long)o * sz + col;
            pw[0] = irho[i];
            pw[1] = m1[i];
            pw[2] = m2[i];
            pw[3] = m3[i];
            pw[4] = m4[i];
            pw[5] = rx[i];
            pw[6] = ry[i];
            pw[7] = rz[i];
            pw[8] = u[i];
            pw[9] = w[i];
        }
    };

    float qzx1[9], qyz1[9], qzx2[9], qyz2[9]; // mixed partial sums: Q[j] <-> output plane p - 4 + j
    float xy1[5], xy2[5];                     // F_xy delayed: xy[4] = F_xy(p)
    float dz1[8], dz2[8];                     // dz[k] <-> d_z(p - 7 + k)
#pragma unroll
    for (int j = 0; j < 9; ++j)
        qzx1[j] = qyz1[j] = qzx2[j] = qyz2[j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 5; ++j)
        xy1[j] = xy2[j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        dz1[j] = dz2[j] = 0.0f;
    float n1z = 0.0f, n2z = 0.0f; // d_z of the next plane
    if (active) {
        n1z = d1z[(long)p0 * sz + col];
        n2z = d2z[(long)p0 * sz + col];
    }
    load_pw(p0 - R + 1);
    fetch(p0);
    commit(p0, 0);
    if (p0 + 1 < pend)
        fetch(p0 + 1);
    __syncthreads();

    for (int p = p0; p < pend; ++p) {
        const int o = p - R, b = (p - p0) & 1;
        const bool emit = o >= z0;
        dz1[7] = n1z;
        dz2[7] = n2z;
        if (active && p + 1 < pend) {
            n1z = d1z[(long)(p + 1) * sz + col];
            n2z = d2z[(long)(p + 1) * sz + col];
        }
        if (p + 1 < pend)
            commit(p + 1, b ^ 1);
        if (p + 2 < pend)
            fetch(p + 2);
        float un = 0.0f, wn = 0.0f;
        if (active) {
            const float(*S1)[W] = s1[b];
            const float(*S2)[W] = s2[b];
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int k = R - j;
                if (k != 0) {
                    const int ak = k < 0 ? -k : k;
                    const float e = ecen(ak);
                    qzx1[j] += e * (S1[ty + R][tx + R + k] - S1[ty + R][tx + R - k]);
                    qyz1[j] += e * (S1[ty + R + k][tx + R] - S1[ty + R - k][tx + R]);
                    qzx2[j] += e * (S2[ty + R][tx + R + k] - S2[ty + R][tx + R - k]);
                    qyz2[j] += e * (S2[ty + R + k][tx + R] - S2[ty + R - k][tx + R]);
                }
            }
            float a1 = 0, a2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float e = ecen(n);
                a1 += e * ((S1[ty + R + n][tx + R + n] - S1[ty + R - n][tx + R + n]) +
                           (S1[ty + R - n][tx + R - n] - S1[ty + R + n][tx + R - n]));
                a2 += e * ((S2[ty + R + n][tx + R + n] - S2[ty + R - n][tx + R + n]) +
                           (S2[ty + R - n][tx + R - n] - S2[ty + R + n][tx + R - n]));
            }
            xy1[4] = a1 * invdxdy;
            xy2[4] = a2 * invdxdy;

            if (emit) {
                const float(*X1)[W] = x1[b];
                const float(*X2)[W] = x2[b];
                const float(*Y1)[BX] = y1[b];
                const float(*Y2)[BX] = y2[b];
                float fxx1 = 0, fxx2 = 0, fyy1 = 0, fyy2 = 0, fzz1 = 0, fzz2 = 0;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    fxx1 += c * (X1[ty][tx + R + n] - X1[ty][tx + R - n + 1]);
                    fxx2 += c * (X2[ty][tx + R + n] - X2[ty][tx + R - n + 1]);
                    fyy1 += c * (Y1[ty + R + n][tx] - Y1[ty + R - n + 1][tx]);
                    fyy2 += c * (Y2[ty + R + n][tx] - Y2[ty + R - n + 1][tx]);
                    fzz1 += c * (dz1[3 + n] - dz1[4 - n]);
                    fzz2 += c * (dz2[3 + n] - dz2[4 - n]);
                }
                const float rho = 1.0f / pw[0];
                fxx1 *= invdx * rho;
                fxx2 *= invdx * rho;
                fyy1 *= invdy * rho;
                fyy2 *= invdy * rho;
                fzz1 *= invdz * rho;
                fzz2 *= invdz * rho;
                const float fzx1 = qzx1[0] * invdzdx, fzx2 = qzx2[0] * invdzdx;
                const float fyz1 = qyz1[0] * invdydz, fyz2 = qyz2[0] * invdydz;
                const float fxy1 = xy1[0], fxy2 = xy2[0];
                const float vx = pw[5], vy = pw[6], vz = pw[7];
                const float rx2 = vx * vx, ry2 = vy * vy, rz2 = vz * vz;
                const float rxy = 2.0f * vx * vy, ryz = 2.0f * vy * vz, rzx = 2.0f * vz * vx;
                const float V1 = rx2 * fxx1 + ry2 * fyy1 + rz2 * fzz1 + rxy * fxy1 + rzx * fzx1 + ryz * fyz1;
                const float H1 = fxx1 + fyy1 + fzz1 - V1;
                const float V2 = rx2 * fxx2 + ry2 * fyy2 + rz2 * fzz2 + rxy * fxy2 + rzx * fzx2 + ryz * fyz2;
                const float H2 = fxx2 + fyy2 + fzz2 - V2;
                const float mm1 = pw[1], mm2 = pw[2], mm3 = pw[3], mm4 = pw[4];
                const float r3 = mm1 * H1 + mm4 * V1 + mm2 * V2;
                const float r4 = mm2 * H1 + mm3 * V2 + mm4 * H2;
                un = pw[8] + r3 * dt;
                wn = pw[9] + r4 * dt;
            }
        }
        // the pointwise words of the next output plane (their registers are free now), then this plane's stores
        load_pw(o + 1);
        if (active && emit) {
            const long i = (long)o * sz + col;
            u[i] = un;
            w[i] = wn;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            qzx1[j] = qzx1[j + 1];
            qyz1[j] = qyz1[j + 1];
            qzx2[j] = qzx2[j + 1];
            qyz2[j] = qyz2[j + 1];
        }
        qzx1[8] = qyz1[8] = qzx2[8] = qyz2[8] = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            xy1[j] = xy1[j + 1];
            xy2[j] = xy2[j + 1];
        }
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            dz1[j] = dz1[j + 1];
            dz2[j] = dz2[j + 1];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// a6: adjoint derivatives (kernels/kernel_type2_grad_snap.cu:242-427; launcher :429-477)
//   S_{f,a} = rho (lap_f + r_a^2 hz_f), M_{f,ab} = 2 r_a r_b hz_f at every tile point;
//   d_{f,a} = b D-_a S_{f,a};  rhs_f = X_xy M_xy + X_yz M_yz + X_zx M_zx.
// ------------------------------------------------------------------------------------------------------------------
struct AdjTemps
{
    float lap1, hz1, lap2, hz2;
};

__device__ __forceinline__ AdjTemps
adj_temps(float f1, float f2, float m1, float m2, float m3, float m4)
{
    AdjTemps t;
    t.lap1 = m1 * f1 + m2 * f2;
    t.hz1 = (-m1 + m4) * f1 - m2 * f2;
    t.lap2 = m4 * f2;
    t.hz2 = m2 * f1 + (-m4 + m3) * f2;
    return t;
}

template <int BY>
__global__ __launch_bounds__(BX* BY) void k_adj_drv(
    float* __restrict__ d1x, float* __restrict__ d2x, float* __restrict__ d1y, float* __restrict__ d2y,
    float* __restrict__ d1z, float* __restrict__ d2z, float* __restrict__ rhs3, float* __restrict__ rhs4,
    const float* __restrict__ f1, const float* __restrict__ f2, const float* __restrict__ b,
    const float* __restrict__ m1, const float* __restrict__ m2, const float* __restrict__ m3,
    const float* __restrict__ m4, const float* __restrict__ m5, const float* __restrict__ m6,
    const float* __restrict__ m7, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx,
    long ldy, float invdx, float invdy, float invdz, float invdxy, float invdyz, float invdzx, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R;
    __shared__ float sx1[BY][W], sx2[BY][W];   // S_x (x halo)
    __shared__ float sy1[H][BX], sy2[H][BX];   // S_y (y halo)
    __shared__ float mxy1[H][W], mxy2[H][W];   // M_xy (both halos)
    __shared__ float mzx1[BY][W], mzx2[BY][W]; // M_zx (x halo)
    __shared__ float myz1[H][BX], myz2[H][BX]; // M_yz (y halo)
    __shared__ float bt[BY][BX];               // b at plane p (no halo)
    const int tx = threadIdx.x, ty = threadIdx.y;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi);
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const long col = (long)iy * ldx + ix;

    float qr1[9], qr2[9];                     // partial X_yz + X_zx sums, q[j] <-> output p - 4 + j
    float dx1[5], dx2[5], dy1[5], dy2[5];     // delayed d_x, d_y (unscaled by b) at planes p-4..p
    float rxy1[5], rxy2[5];                   // delayed X_xy
    float szq1[9], szq2[9];                   // S_z(p - 8 + k)
    float bq[5];                              // b(p - 4 + k)
#pragma unroll
    for (int j = 0; j < 9; ++j)
        qr1[j] = qr2[j] = szq1[j] = szq2[j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 5; ++j)
        dx1[j] = dx2[j] = dy1[j] = dy2[j] = rxy1[j] = rxy2[j] = bq[j] = 0.0f;

    // The output o = p - 4 needs S_z(o - 4 .. o + 3), absorbed at planes z0 - 4 onwards.
    for (int p = z0 - R; p < z1 + R; ++p) {
        const int o = p - R;
        {
            // temporaries over the (H x W) tile of plane p
            for (int k = ty * BX + tx; k < W * H; k += BX * BY) {
                const int ly = k / W;
                const int lx = k - ly * W;
                float vs1x = 0, vs2x = 0, vs1y = 0, vs2y = 0, vm1xy = 0, vm2xy = 0, vm1zx = 0, vm2zx = 0, vm1yz = 0,
                      vm2yz = 0;
                if (lx < xlim && ly < ylim) {
                    const long g = (long)p * sz + (long)(y0 - R + ly) * ldx + (x0 - R + lx);
                    const AdjTemps t = adj_temps(f1[g], f2[g], m1[g], m2[g], m3[g], m4[g]);
                    const float rho = 1.0f / b[g];
                    const float vx = m5[g], vy = m6[g], vz = m7[g];
                    vs1x = rho * (t.lap1 + vx * vx * t.hz1);
                    vs2x = rho * (t.lap2 + vx * vx * t.hz2);
                    vs1y = rho * (t.lap1 + vy * vy * t.hz1);
                    vs2y = rho * (t.lap2 + vy * vy * t.hz2);
                    vm1xy = 2.0f * vx * vy * t.hz1;
                    vm2xy = 2.0f * vx * vy * t.hz2;
                    vm1zx = 2.0f * vz * vx * t.hz1;
                    vm2zx = 2.0f * vz * vx * t.hz2;
                    vm1yz = 2.0f * vy * vz * t.hz1;
                    vm2yz = 2.0f * vy * vz * t.hz2;
                }
                mxy1[ly][lx] = vm1xy;
                mxy2[ly][lx] = vm2xy;
                if (ly >= R && ly < R + BY) {
                    sx1[ly - R][lx] = vs1x;
                    sx2[ly - R][lx] = vs2x;
                    mzx1[ly - R][lx] = vm1zx;
                    mzx2[ly - R][lx] = vm2zx;
                }
                if (lx >= R && lx < R + BX) {
                    sy1[ly][lx - R] = vs1y;
                    sy2[ly][lx - R] = vs2y;
                    myz1[ly][lx - R] = vm1yz;
                    myz2[ly][lx - R] = vm2yz;
                }
            }
        }
        // own column S_z(p) and b(p)
        if (active) {
            const long g = (long)p * sz + col;
            const AdjTemps t = adj_temps(f1[g], f2[g], m1[g], m2[g], m3[g], m4[g]);
            const float bb = b[g];
            const float rho = 1.0f / bb;
            const float vz = m7[g];
            szq1[8] = rho * (t.lap1 + vz * vz * t.hz1);
            szq2[8] = rho * (t.lap2 + vz * vz * t.hz2);
            bq[4] = bb;
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int k = R - j;
                if (k != 0) {
                    const int ak = k < 0 ? -k : k;
                    const float e = ecen(ak);
                    qr1[j] += invdyz * e * (myz1[ty + R + k][tx] - myz1[ty + R - k][tx]) +
                              invdzx * e * (mzx1[ty][tx + R + k] - mzx1[ty][tx + R - k]);
                    qr2[j] += invdyz * e * (myz2[ty + R + k][tx] - myz2[ty + R - k][tx]) +
                              invdzx * e * (mzx2[ty][tx + R + k] - mzx2[ty][tx + R - k]);
                }
            }
            float ax1 = 0, ax2 = 0, ay1 = 0, ay2 = 0, axy1 = 0, axy2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                ax1 += c * (sx1[ty][tx + R + n - 1] - sx1[ty][tx + R - n]);
                ax2 += c * (sx2[ty][tx + R + n - 1] - sx2[ty][tx + R - n]);
                ay1 += c * (sy1[ty + R + n - 1][tx] - sy1[ty + R - n][tx]);
                ay2 += c * (sy2[ty + R + n - 1][tx] - sy2[ty + R - n][tx]);
                const float e = ecen(n);
                axy1 += e * (mxy1[ty + R + n][tx + R + n] - mxy1[ty + R + n][tx + R - n] +
                             mxy1[ty + R - n][tx + R - n] - mxy1[ty + R - n][tx + R + n]);
                axy2 += e * (mxy2[ty + R + n][tx + R + n] - mxy2[ty + R + n][tx + R - n] +
                             mxy2[ty + R - n][tx + R - n] - mxy2[ty + R - n][tx + R + n]);
            }
            dx1[4] = ax1 * invdx;
            dx2[4] = ax2 * invdx;
            dy1[4] = ay1 * invdy;
            dy2[4] = ay2 * invdy;
            rxy1[4] = axy1 * invdxy;
            rxy2[4] = axy2 * invdxy;
            if (o >= z0) {
                float az1 = 0, az2 = 0;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    // S_z index: szq[k] <-> p - 8 + k, output o = p - 4 -> o + n - 1 = k of (3 + n), o - n = (4 - n)
                    az1 += c * (szq1[3 + n] - szq1[4 - n]);
                    az2 += c * (szq2[3 + n] - szq2[4 - n]);
                }
                const float bo = bq[0];
                const long i = (long)o * sz + col;
                d1x[i] = dx1[0] * bo;
                d2x[i] = dx2[0] * bo;
                d1y[i] = dy1[0] * bo;
                d2y[i] = dy2[0] * bo;
                d1z[i] = az1 * invdz * bo;
                d2z[i] = az2 * invdz * bo;
                rhs3[i] = rxy1[0] + qr1[0];
                rhs4[i] = rxy2[0] + qr2[0];
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            qr1[j] = qr1[j + 1];
            qr2[j] = qr2[j + 1];
            szq1[j] = szq1[j + 1];
            szq2[j] = szq2[j + 1];
        }
        qr1[8] = qr2[8] = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            dx1[j] = dx1[j + 1];
            dx2[j] = dx2[j + 1];
            dy1[j] = dy1[j + 1];
            dy2[j] = dy2[j + 1];
            rxy1[j] = rxy1[j + 1];
            rxy2[j] = rxy2[j + 1];
            bq[j] = bq[j + 1];
        }
    }
}

// The same operator with the plane pipeline: the 10 input fields of the (H x W) tile of plane p + 2 are loaded as
// float4 quads (one quad per thread) during plane p; during plane p + 1 the quads become the temporaries S_x, S_y,
// M_xy, M_zx, M_yz (and, on the centre, S_z and b of every column) in the other LDS buffer; one barrier per plane.
// Needs quads_fit (launcher falls back).
template <int BY>
__global__ __launch_bounds__(BX* BY) void k_adj_drv_q(
    float* __restrict__ d1x, float* __restrict__ d2x, float* __restrict__ d1y, float* __restrict__ d2y,
    float* __restrict__ d1z, float* __restrict__ d2z, float* __restrict__ rhs3, float* __restrict__ rhs4,
    const float* __restrict__ f1, const float* __restrict__ f2, const float* __restrict__ b,
    const float* __restrict__ m1, const float* __restrict__ m2, const float* __restrict__ m3,
    const float* __restrict__ m4, const float* __restrict__ m5, const float* __restrict__ m6,
    const float* __restrict__ m7, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx,
    long ldy, float invdx, float invdy, float invdz, float invdxy, float invdyz, float invdzx, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R, QPR = W / 4, NQ = QPR * H;
    static_assert(NQ <= BX * BY, "one tile quad per thread");
    __shared__ __align__(16) float sx1[2][BY][W], sx2[2][BY][W];   // S_x (x halo)
    __shared__ __align__(16) float sy1[2][H][BX], sy2[2][H][BX];   // S_y (y halo)
    __shared__ __align__(16) float mxy1[2][H][W], mxy2[2][H][W];   // M_xy (both halos)
    __shared__ __align__(16) float mzx1[2][BY][W], mzx2[2][BY][W]; // M_zx (x halo)
    __shared__ __align__(16) float myz1[2][H][BX], myz2[2][H][BX]; // M_yz (y halo)
    __shared__ __align__(16) float bt[2][BY][BX];                  // b (centre)
    __shared__ __align__(16) float szt1[2][BY][BX], szt2[2][BY][BX]; // S_z (centre)
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi);
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const long col = (long)iy * ldx + ix;
    const int p0 = z0 - R, pend = z1 + R; // tile planes p in [p0, pend), output o = p - R

    // this thread's tile quad: row qr, columns ql .. ql + 3 (elements at or beyond xlim / ylim are zero)
    const bool qhas = tid < NQ;
    const int qr = tid / QPR, ql = 4 * (tid - qr * QPR);
    const int qkeep = qhas && qr < ylim ? min(max(xlim - ql, 0), 4) : 0;
    const long qoff = (long)(y0 - R + qr) * ldx + (x0 - R + ql);
    const bool qrow = qr >= R && qr < R + BY, qcol = ql >= R && ql < R + BX;
    float4 v1, v2, vb, vm1, vm2, vm3, vm4, vm5, vm6, vm7;
    auto fetch = [&](int p) {
        if (qkeep > 0) {
            const long g = (long)p * sz + qoff;
            v1 = *reinterpret_cast<const float4*>(f1 + g);
            v2 = *reinterpret_cast<const float4*>(f2 + g);
            vb = *reinterpret_cast<const float4*>(b + g);
            vm1 = *reinterpret_cast<const float4*>(m1 + g);
            vm2 = *reinterpret_cast<const float4*>(m2 + g);
            vm3 = *reinterpret_cast<const float4*>(m3 + g);
            vm4 = *reinterpret_cast<const float4*>(m4 + g);
            vm5 = *reinterpret_cast<const float4*>(m5 + g);
            vm6 = *reinterpret_cast<const float4*>(m6 + g);
            vm7 = *reinterpret_cast<const float4*>(m7 + g);
        }
    };
    auto commit = [&](int bf) {
        if (!qhas)
            return;
        const float* A1 = &v1.x;
        const float* A2 = &v2.x;
        const float* BB = &vb.x;
        const float* M1 = &vm1.x;
        const float* M2 = &vm2.x;
        const float* M3 = &vm3.x;
        const float* M4 = &vm4.x;
        const float* M5 = &vm5.x;
        const float* M6 = &vm6.x;
        const float* M7 = &vm7.x;
        float s1x[4], s2x[4], s1y[4], s2y[4], a1xy[4], a2xy[4], a1zx[4], a2zx[4], a1yz[4], a2yz[4], s1z[4], s2z[4],
            bv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            s1x[e] = s2x[e] = s1y[e] = s2y[e] = a1xy[e] = a2xy[e] = 0.0f;
            a1zx[e] = a2zx[e] = a1yz[e] = a2yz[e] = s1z[e] = s2z[e] = bv[e] = 0.0f;
            if (e < qkeep) {
                const AdjTemps t = adj_temps(A1[e], A2[e], M1[e], M2[e], M3[e], M4[e]);
                const float rho = 1.0f / BB[e];
                const float vx = M5[e], vy = M6[e], vz = M7[e];
                s1x[e] = rho * (t.lap1 + vx * vx * t.hz1);
                s2x[e] = rho * (t.lap2 + vx * vx * t.hz2);
                s1y[e] = rho * (t.lap1 + vy * vy * t.hz1);
                s2y[e] = rho * (t.lap2 + vy * vy * t.hz2);
                a1xy[e] = 2.0f * vx * vy * t.hz1;
                a2xy[e] = 2.0f * vx * vy * t.hz2;
                a1zx[e] = 2.0f * vz * vx * t.hz1;
                a2zx[e] = 2.0f * vz * vx * t.hz2;
                a1yz[e] = 2.0f * vy * vz * t.hz1;
                a2yz[e] = 2.0f * vy * vz * t.hz2;
                s1z[e] = rho * (t.lap1 + vz * vz * t.hz1);
                s2z[e] = rho * (t.lap2 + vz * vz * t.hz2);
                bv[e] = BB[e];
            }
        }
        auto st = [](float* dst, const float* v) {
            *reinterpret_cast<float4*>(dst) = make_float4(v[0], v[1], v[2], v[3]);
        };
        st(&mxy1[bf][qr][ql], a1xy);
        st(&mxy2[bf][qr][ql], a2xy);
        if (qrow) {
            st(&sx1[bf][qr - R][ql], s1x);
            st(&sx2[bf][qr - R][ql], s2x);
            st(&mzx1[bf][qr - R][ql], a1zx);
            st(&mzx2[bf][qr - R][ql], a2zx);
        }
        if (qcol) {
            st(&sy1[bf][qr][ql - R], s1y);
            st(&sy2[bf][qr][ql - R], s2y);
            st(&myz1[bf][qr][ql - R], a1yz);
            st(&myz2[bf][qr][ql - R], a2yz);
        }
        if (qrow && qcol) {
            st(&bt[bf][qr - R][ql - R], bv);
            st(&szt1[bf][qr - R][ql - R], s1z);
            st(&szt2[bf][qr - R][ql - R], s2z);
        }
    };

    float qr1[9], qr2[9];                 // partial X_yz + X_zx sums, q[j] <-> output p - 4 + j
    float dx1[5], dx2[5], dy1[5], dy2[5]; // delayed d_x, d_y (unscaled by b) at planes p-4..p
    float rxy1[5], rxy2[5];               // delayed X_xy
    float szq1[9], szq2[9];               // S_z(p - 8 + k)
    float bq[5];                          // b(p - 4 + k)
#pragma unroll
    for (int j = 0; j < 9; ++j)
        qr1[j] = qr2[j] = szq1[j] = szq2[j] = 0.0f;
#pragma unroll
    for (int j = 0; j < 5; ++j)
        dx1[j] = dx2[j] = dy1[j] = dy2[j] = rxy1[j] = rxy2[j] = bq[j] = 0.0f;

    fetch(p0);
    commit(0);
    if (p0 + 1 < pend)
        fetch(p0 + 1);
    __syncthreads();
    // The output o = p - 4 needs S_z(o - 4 .. o + 3), absorbed at planes z0 - 4 onwards.
    for (int p = p0; p < pend; ++p) {
        const int o = p - R, bf = (p - p0) & 1;
        if (p + 1 < pend)
            commit(bf ^ 1);
        if (p + 2 < pend)
            fetch(p + 2);
        if (active) {
            szq1[8] = szt1[bf][ty][tx];
            szq2[8] = szt2[bf][ty][tx];
            bq[4] = bt[bf][ty][tx];
            const float(*SX1)[W] = sx1[bf];
            const float(*SX2)[W] = sx2[bf];
            const float(*SY1)[BX] = sy1[bf];
            const float(*SY2)[BX] = sy2[bf];
            const float(*MXY1)[W] = mxy1[bf];
            const float(*MXY2)[W] = mxy2[bf];
            const float(*MZX1)[W] = mzx1[bf];
            const float(*MZX2)[W] = mzx2[bf];
            const float(*MYZ1)[BX] = myz1[bf];
            const float(*MYZ2)[BX] = myz2[bf];
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const int k = R - j;
                if (k != 0) {
                    const int ak = k < 0 ? -k : k;
                    const float e = ecen(ak);
                    qr1[j] += invdyz * e * (MYZ1[ty + R + k][tx] - MYZ1[ty + R - k][tx]) +
                              invdzx * e * (MZX1[ty][tx + R + k] - MZX1[ty][tx + R - k]);
                    qr2[j] += invdyz * e * (MYZ2[ty + R + k][tx] - MYZ2[ty + R - k][tx]) +
                              invdzx * e * (MZX2[ty][tx + R + k] - MZX2[ty][tx + R - k]);
                }
            }
            float ax1 = 0, ax2 = 0, ay1 = 0, ay2 = 0, axy1 = 0, axy2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                ax1 += c * (SX1[ty][tx + R + n - 1] - SX1[ty][tx + R - n]);
                ax2 += c * (SX2[ty][tx + R + n - 1] - SX2[ty][tx + R - n]);
                ay1 += c * (SY1[ty + R + n - 1][tx] - SY1[ty + R - n][tx]);
                ay2 += c * (SY2[ty + R + n - 1][tx] - SY2[ty + R - n][tx]);
                const float e = ecen(n);
                axy1 += e * (MXY1[ty + R + n][tx + R + n] - MXY1[ty + R + n][tx + R - n] +
                             MXY1[ty + R - n][tx + R - n] - MXY1[ty + R - n][tx + R + n]);
                axy2 += e * (MXY2[ty + R + n][tx + R + n] - MXY2[ty + R + n][tx + R - n] +
                             MXY2[ty + R - n][tx + R - n] - MXY2[ty + R - n][tx + R + n]);
            }
            dx1[4] = ax1 * invdx;
            dx2[4] = ax2 * invdx;
            dy1[4] = ay1 * invdy;
            dy2[4] = ay2 * invdy;
            rxy1[4] = axy1 * invdxy;
            rxy2[4] = axy2 * invdxy;
            if (o >= z0) {
                float az1 = 0, az2 = 0;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    az1 += c * (szq1[3 + n] - szq1[4 - n]);
                    az2 += c * (szq2[3 + n] - szq2[4 - n]);
                }
                const float bo = bq[0];
                const long i = (long)o * sz + col;
                d1x[i] = dx1[0] * bo;
                d2x[i] = dx2[0] * bo;
                d1y[i] = dy1[0] * bo;
                d2y[i] = dy2[0] * bo;
                d1z[i] = az1 * invdz * bo;
                d2z[i] = az2 * invdz * bo;
                rhs3[i] = rxy1[0] + qr1[0];
                rhs4[i] = rxy2[0] + qr2[0];
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            qr1[j] = qr1[j + 1];
            qr2[j] = qr2[j + 1];
            szq1[j] = szq1[j + 1];
            szq2[j] = szq2[j + 1];
        }
        qr1[8] = qr2[8] = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            dx1[j] = dx1[j + 1];
            dx2[j] = dx2[j + 1];
            dy1[j] = dy1[j + 1];
            dy2[j] = dy2[j + 1];
            rxy1[j] = rxy1[j + 1];
            rxy2[j] = rxy2[j + 1];
            bq[j] = bq[j + 1];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// a7: adjoint update (kernels/kernel_type2_adj.cu:100-260; launchers :262-349)
//   r = sum_a D+_a d_a + rhs_mixed;  u += dt r   or, with the sponge, u = A u + B dt r.
// ------------------------------------------------------------------------------------------------------------------
template <int BY, bool SPONGE>
__global__ __launch_bounds__(BX* BY) void k_adj_update(
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ d1x, const float* __restrict__ d2x,
    const float* __restrict__ d1y, const float* __restrict__ d2y, const float* __restrict__ d1z,
    const float* __restrict__ d2z, const float* __restrict__ rhs3, const float* __restrict__ rhs4,
    const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, int ixbeg, int ixend,
    int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy, float dt, float invdx, float invdy, float invdz,
    int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R;
    __shared__ float x1[BY][W], x2[BY][W];
    __shared__ float y1[H][BX], y2[H][BX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    int bxi, byi, bzi;
    bxi = blockIdx.x; // dispatch order: measured faster here than the XCD order
    byi = blockIdx.y;
    bzi = blockIdx.z;
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const int xlim_c = min(BX, ixend + 1 - x0);
    const int ylim_c = min(BY, iyend + 1 - y0);
    const long col = (long)iy * ldx + ix;
    float axy = 1.0f, bxy = 1.0f;
    if (SPONGE && active) {
        axy = fminf(abx[ix].a, aby[iy].a);
        bxy = fminf(abx[ix].b, aby[iy].b);
    }

    float dz1[8], dz2[8]; // d_z(o - 3 + k)
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        dz1[k] = active ? d1z[(long)(z0 - 3 + k) * sz + col] : 0.0f;
        dz2[k] = active ? d2z[(long)(z0 - 3 + k) * sz + col] : 0.0f;
    }
    for (int o = z0; o < z1; ++o) {
        if (active) {
            dz1[7] = d1z[(long)(o + 4) * sz + col];
            dz2[7] = d2z[(long)(o + 4) * sz + col];
        }
        const long orgx = (long)o * sz + (long)y0 * ldx + (x0 - R);
        load_tile<BY, W, BY>(x1, d1x, orgx, ldx, xlim, ylim_c);
        load_tile<BY, W, BY>(x2, d2x, orgx, ldx, xlim, ylim_c);
        const long orgy = (long)o * sz + (long)(y0 - R) * ldx + x0;
        load_tile<BY, BX, H>(y1, d1y, orgy, ldx, xlim_c, ylim);
        load_tile<BY, BX, H>(y2, d2y, orgy, ldx, xlim_c, ylim);
        __syncthreads();
        if (active) {
            float fx1 = 0, fx2 = 0, fy1 = 0, fy2 = 0, fz1 = 0, fz2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                fx1 += c * (x1[ty][tx + R + n] - x1[ty][tx + R - n + 1]);
                fx2 += c * (x2[ty][tx + R + n] - x2[ty][tx + R - n + 1]);
                fy1 += c * (y1[ty + R + n][tx] - y1[ty + R - n + 1][tx]);
                fy2 += c * (y2[ty + R + n][tx] - y2[ty + R - n + 1][tx]);
                fz1 += c * (dz1[3 + n] - dz1[4 - n]);
                fz2 += c * (dz2[3 + n] - dz2[4 - n]);
            }
            const long i = (long)o * sz + col;
            const float r3 = fx1 * invdx + fy1 * invdy + fz1 * invdz + rhs3[i];
            const float r4 = fx2 * invdx + fy2 * invdy + fz2 * invdz + rhs4[i];
            if (SPONGE) {
                const float a = fminf(axy, abz[o].a);
                const float bb = fminf(bxy, abz[o].b) * dt;
                u[i] = a * u[i] + bb * r3;
                w[i] = a * w[i] + bb * r4;
            } else {
                u[i] += r3 * dt;
                w[i] += r4 * dt;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            dz1[k] = dz1[k + 1];
            dz2[k] = dz2[k + 1];
        }
    }
}

// The same operator with the plane pipeline: the d_x / d_y tiles of plane o + 1 are loaded as float4 quads during
// plane o - 1 and stored into the other LDS buffer during plane o; the d_z column word and the pointwise words
// (rhs3, rhs4, u, w) of plane o + 1 are loaded during plane o; one barrier per plane. Needs quads_fit.
template <int BY, bool SPONGE>
__global__ __launch_bounds__(BX* BY) void k_adj_update_q(
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ d1x, const float* __restrict__ d2x,
    const float* __restrict__ d1y, const float* __restrict__ d2y, const float* __restrict__ d1z,
    const float* __restrict__ d2z, const float* __restrict__ rhs3, const float* __restrict__ rhs4,
    const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, int ixbeg, int ixend,
    int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy, float dt, float invdx, float invdy, float invdz,
    int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R, NT = BX * BY;
    __shared__ __align__(16) float x1[2][BY][W], x2[2][BY][W];
    __shared__ __align__(16) float y1[2][H][BX], y2[2][H][BX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    int bxi, byi, bzi;
    bxi = blockIdx.x; // dispatch order: measured faster here than the XCD order
    byi = blockIdx.y;
    bzi = blockIdx.z;
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const int xlim_c = min(BX, ixend + 1 - x0);
    const int ylim_c = min(BY, iyend + 1 - y0);
    const long col = (long)iy * ldx + ix;
    float axy = 1.0f, bxy = 1.0f;
    if (SPONGE && active) {
        axy = fminf(abx[ix].a, aby[iy].a);
        bxy = fminf(abx[ix].b, aby[iy].b);
    }
    const TileQuads<W, BY, NT> qx(tid, ldx, xlim, ylim_c);
    const TileQuads<BX, H, NT> qy(tid, ldx, xlim_c, ylim);
    const long ox = (long)y0 * ldx + (x0 - R), oy = (long)(y0 - R) * ldx + x0;
    float4 rx1[qx.QT], rx2[qx.QT], ry1[qy.QT], ry2[qy.QT];
    auto fetch = [&](int o) {
        const long oz = (long)o * sz;
        qx.fetch(rx1, d1x + oz + ox);
        qx.fetch(rx2, d2x + oz + ox);
        qy.fetch(ry1, d1y + oz + oy);
        qy.fetch(ry2, d2y + oz + oy);
    };
    auto commit = [&](int bf) {
        qx.commit(&x1[bf][0][0], rx1);
        qx.commit(&x2[bf][0][0], rx2);
        qy.commit(&y1[bf][0][0], ry1);
        qy.commit(&y2[bf][0][0], ry2);
    };
    float pw[4]; // rhs3, rhs4, u, w of the plane being finalised
    auto load_pw = [&](int o) {
        if (active && o < z1) {
            const long i = (long)o * sz + col;
            pw[0] = rhs3[i];
            pw[1] = rhs4[i];
            pw[2] = u[i];
            pw[3] = w[i];
        }
    };

    float dz1[8], dz2[8]; // d_z(o - 3 + k)
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        dz1[k] = active ? d1z[(long)(z0 - 3 + k) * sz + col] : 0.0f;
        dz2[k] = active ? d2z[(long)(z0 - 3 + k) * sz + col] : 0.0f;
    }
    float n1 = 0.0f, n2 = 0.0f; // d_z(o + 4) of the next plane
    if (active) {
        n1 = d1z[(long)(z0 + 4) * sz + col];
        n2 = d2z[(long)(z0 + 4) * sz + col];
    }
    load_pw(z0);
    fetch(z0);
    commit(0);
    if (z0 + 1 < z1)
        fetch(z0 + 1);
    __syncthreads();
    for (int o = z0; o < z1; ++o) {
        const int bf = (o - z0) & 1;
        dz1[7] = n1;
        dz2[7] = n2;
        if (active && o + 1 < z1) {
            n1 = d1z[(long)(o + 5) * sz + col];
            n2 = d2z[(long)(o + 5) * sz + col];
        }
        if (o + 1 < z1)
            commit(bf ^ 1);
        if (o + 2 < z1)
            fetch(o + 2);
        float un = 0.0f, wn = 0.0f;
        if (active) {
            const float(*X1)[W] = x1[bf];
            const float(*X2)[W] = x2[bf];
            const float(*Y1)[BX] = y1[bf];
            const float(*Y2)[BX] = y2[bf];
            float fx1 = 0, fx2 = 0, fy1 = 0, fy2 = 0, fz1 = 0, fz2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                fx1 += c * (X1[ty][tx + R + n] - X1[ty][tx + R - n + 1]);
                fx2 += c * (X2[ty][tx + R + n] - X2[ty][tx + R - n + 1]);
                fy1 += c * (Y1[ty + R + n][tx] - Y1[ty + R - n + 1][tx]);
                fy2 += c * (Y2[ty + R + n][tx] - Y2[ty + R - n + 1][tx]);
                fz1 += c * (dz1[3 + n] - dz1[4 - n]);
                fz2 += c * (dz2[3 + n] - dz2[4 - n]);
            }
            const float r3 = fx1 * invdx + fy1 * invdy + fz1 * invdz + pw[0];
            const float r4 = fx2 * invdx + fy2 * invdy + fz2 * invdz + pw[1];
            if (SPONGE) {
                const float a = fminf(axy, abz[o].a);
                const float bb = fminf(bxy, abz[o].b) * dt;
                un = a * pw[2] + bb * r3;
                wn = a * pw[3] + bb * r4;
            } else {
                un = pw[2] + r3 * dt;
                wn = pw[3] + r4 * dt;
            }
        }
        load_pw(o + 1);
        if (active) {
            const long i = (long)o * sz + col;
            u[i] = un;
            w[i] = wn;
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            dz1[k] = dz1[k + 1];
            dz2[k] = dz2[k + 1];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// a10: density gradient (kernels/kernel_type2_grad.cu:344-509 optimised, :511-560 simple; launcher :565-672)
//   g += -b * sum_{f,a} [ S_{f,a} D+_a Dfwd_{f,a} + (D-_a S_{f,a}) Dfwd_{f,a} ]
//   S from the adjoint sigma; the gradient pointer is not shifted (g index relative to the interior origin).
// ------------------------------------------------------------------------------------------------------------------
template <int BY, class GT = float>
__global__ __launch_bounds__(BX* BY) void k_rho_grad(
    GT* __restrict__ grad, const float* __restrict__ fd1x, const float* __restrict__ fd2x,
    const float* __restrict__ fd1y, const float* __restrict__ fd2y, const float* __restrict__ fd1z,
    const float* __restrict__ fd2z, const float* __restrict__ f1, const float* __restrict__ f2,
    const float* __restrict__ b, const float* __restrict__ m1, const float* __restrict__ m2,
    const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ vrx,
    const float* __restrict__ vry, const float* __restrict__ vrz, int ixbeg, int ixend, int iybeg, int iyend,
    int izbeg, int izend, long ldx, long ldy, long gldx, long gldy, float hx, float hy, float hz, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R;
    __shared__ float sx1[BY][W], sx2[BY][W];
    __shared__ float sy1[H][BX], sy2[H][BX];
    __shared__ float fx1[BY][W], fx2[BY][W];
    __shared__ float fy1[H][BX], fy2[H][BX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi);
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const int xlim_c = min(BX, ixend + 1 - x0);
    const int ylim_c = min(BY, iyend + 1 - y0);
    const long col = (long)iy * ldx + ix;

    float szq1[9], szq2[9]; // S_z(o - 4 + k), k = 0..8
    float fzq1[8], fzq2[8]; // Dfwd_z(o - 3 + k), k = 0..7
    // prologue: S_z for o-4..o+3 and Dfwd_z for o-3..o+3 of the first output
#pragma unroll
    for (int k = 0; k < 9; ++k)
        szq1[k] = szq2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        fzq1[k] = fzq2[k] = 0.0f;
    if (active) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const long g = (long)(z0 - 4 + k) * sz + col;
            const AdjTemps t = adj_temps(f1[g], f2[g], m1[g], m2[g], m3[g], m4[g]);
            const float rho = 1.0f / b[g];
            const float vz = vrz[g];
            szq1[k] = rho * (t.lap1 + vz * vz * t.hz1);
            szq2[k] = rho * (t.lap2 + vz * vz * t.hz2);
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            fzq1[k] = fd1z[(long)(z0 - 3 + k) * sz + col];
            fzq2[k] = fd2z[(long)(z0 - 3 + k) * sz + col];
        }
    }
    for (int o = z0; o < z1; ++o) {
        // S tiles of plane o
        for (int k = ty * BX + tx; k < W * H; k += BX * BY) {
            const int ly = k / W;
            const int lx = k - ly * W;
            const bool inrow = ly >= R && ly < R + BY;
            const bool incol = lx >= R && lx < R + BX;
            if (!inrow && !incol)
                continue;
            float vs1x = 0, vs2x = 0, vs1y = 0, vs2y = 0;
            if (lx < xlim && ly < ylim) {
                const long g = (long)o * sz + (long)(y0 - R + ly) * ldx + (x0 - R + lx);
                const AdjTemps t = adj_temps(f1[g], f2[g], m1[g], m2[g], m3[g], m4[g]);
                const float rho = 1.0f / b[g];
                const float vx = vrx[g], vy = vry[g];
                vs1x = rho * (t.lap1 + vx * vx * t.hz1);
                vs2x = rho * (t.lap2 + vx * vx * t.hz2);
                vs1y = rho * (t.lap1 + vy * vy * t.hz1);
                vs2y = rho * (t.lap2 + vy * vy * t.hz2);
            }
            if (inrow) {
                sx1[ly - R][lx] = vs1x;
                sx2[ly - R][lx] = vs2x;
            }
            if (incol) {
                sy1[ly][lx - R] = vs1y;
                sy2[ly][lx - R] = vs2y;
            }
        }
        const long orgx = (long)o * sz + (long)y0 * ldx + (x0 - R);
        load_tile<BY, W, BY>(fx1, fd1x, orgx, ldx, xlim, ylim_c);
        load_tile<BY, W, BY>(fx2, fd2x, orgx, ldx, xlim, ylim_c);
        const long orgy = (long)o * sz + (long)(y0 - R) * ldx + x0;
        load_tile<BY, BX, H>(fy1, fd1y, orgy, ldx, xlim_c, ylim);
        load_tile<BY, BX, H>(fy2, fd2y, orgy, ldx, xlim_c, ylim);
        if (active) {
            const long g = (long)(o + 4) * sz + col;
            const AdjTemps t = adj_temps(f1[g], f2[g], m1[g], m2[g], m3[g], m4[g]);
            const float rho = 1.0f / b[g];
            const float vz = vrz[g];
            szq1[8] = rho * (t.lap1 + vz * vz * t.hz1);
            szq2[8] = rho * (t.lap2 + vz * vz * t.hz2);
            fzq1[7] = fd1z[g];
            fzq2[7] = fd2z[g];
        }
        __syncthreads();
        if (active) {
            float dpx1 = 0, dpx2 = 0, dpy1 = 0, dpy2 = 0, dpz1 = 0, dpz2 = 0; // D+ Dfwd
            float dmx1 = 0, dmx2 = 0, dmy1 = 0, dmy2 = 0, dmz1 = 0, dmz2 = 0; // D- S
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                dpx1 += c * (fx1[ty][tx + R + n] - fx1[ty][tx + R - n + 1]);
                dpx2 += c * (fx2[ty][tx + R + n] - fx2[ty][tx + R - n + 1]);
                dpy1 += c * (fy1[ty + R + n][tx] - fy1[ty + R - n + 1][tx]);
                dpy2 += c * (fy2[ty + R + n][tx] - fy2[ty + R - n + 1][tx]);
                dpz1 += c * (fzq1[3 + n] - fzq1[4 - n]);
                dpz2 += c * (fzq2[3 + n] - fzq2[4 - n]);
                dmx1 += c * (sx1[ty][tx + R + n - 1] - sx1[ty][tx + R - n]);
                dmx2 += c * (sx2[ty][tx + R + n - 1] - sx2[ty][tx + R - n]);
                dmy1 += c * (sy1[ty + R + n - 1][tx] - sy1[ty + R - n][tx]);
                dmy2 += c * (sy2[ty + R + n - 1][tx] - sy2[ty + R - n][tx]);
                dmz1 += c * (szq1[4 + n - 1] - szq1[4 - n]);
                dmz2 += c * (szq2[4 + n - 1] - szq2[4 - n]);
            }
            const long i = (long)o * sz + col;
            const float rhs = sx1[ty][tx + R] * (dpx1 * hx) + sx2[ty][tx + R] * (dpx2 * hx) +
                              sy1[ty + R][tx] * (dpy1 * hy) + sy2[ty + R][tx] * (dpy2 * hy) +
                              szq1[4] * (dpz1 * hz) + szq2[4] * (dpz2 * hz) + (dmx1 * hx) * fx1[ty][tx + R] +
                              (dmx2 * hx) * fx2[ty][tx + R] + (dmy1 * hy) * fy1[ty + R][tx] +
                              (dmy2 * hy) * fy2[ty + R][tx] + (dmz1 * hz) * fzq1[3] + (dmz2 * hz) * fzq2[3];
            const long gi = ((long)(o - izbeg) * gldy + (iy - iybeg)) * gldx + (ix - ixbeg);
            grad[gi] += (GT)(-b[i] * rhs);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            szq1[k] = szq1[k + 1];
            szq2[k] = szq2[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            fzq1[k] = fzq1[k + 1];
            fzq2[k] = fzq2[k + 1];
        }
    }
}

// The same operator with the plane pipeline: the inputs of the S_x / S_y tiles (a plus-shaped footprint of float4
// quads, one per thread, 9 fields) and the forward-derivative tiles of plane o + 1 are loaded during plane o - 1 and
// stored (as S_x, S_y, Dfwd_x, Dfwd_y, and b on the centre) into the other LDS buffer during plane o; the column
// words of S_z(o + 5), Dfwd_z(o + 5) and the image word of plane o + 1 are loaded during plane o; one barrier per
// plane. Needs quads_fit (launcher falls back).
template <int BY, class GT = float>
__global__ __launch_bounds__(BX* BY) void k_rho_grad_q(
    GT* __restrict__ grad, const float* __restrict__ fd1x, const float* __restrict__ fd2x,
    const float* __restrict__ fd1y, const float* __restrict__ fd2y, const float* __restrict__ fd1z,
    const float* __restrict__ fd2z, const float* __restrict__ f1, const float* __restrict__ f2,
    const float* __restrict__ b, const float* __restrict__ m1, const float* __restrict__ m2,
    const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ vrx,
    const float* __restrict__ vry, const float* __restrict__ vrz, int ixbeg, int ixend, int iybeg, int iyend,
    int izbeg, int izend, long ldx, long ldy, long gldx, long gldy, float hx, float hy, float hz, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R, NT = BX * BY, QPR = W / 4, NQ = QPR * H;
    static_assert(NQ <= NT, "one tile quad per thread");
    __shared__ __align__(16) float sx1[2][BY][W], sx2[2][BY][W];
    __shared__ __align__(16) float sy1[2][H][BX], sy2[2][H][BX];
    __shared__ __align__(16) float fx1[2][BY][W], fx2[2][BY][W];
    __shared__ __align__(16) float fy1[2][H][BX], fy2[2][H][BX];
    __shared__ __align__(16) float bt[2][BY][BX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    int bxi, byi, bzi;
    stencil_tile(bxi, byi, bzi);
    const int x0 = ixbeg + bxi * BX, y0 = iybeg + byi * BY, z0 = izbeg + bzi * bz;
    const int z1 = min(z0 + bz, izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= ixend && iy <= iyend;
    const long sz = ldx * ldy;
    const int xlim = min(W, ixend + R + 1 - (x0 - R));
    const int ylim = min(H, iyend + R + 1 - (y0 - R));
    const int xlim_c = min(BX, ixend + 1 - x0);
    const int ylim_c = min(BY, iyend + 1 - y0);
    const long col = (long)iy * ldx + ix;

    // this thread's quad of the S footprint (rows of the tile or columns of the tile; corners are not needed)
    const int qr = tid / QPR, ql = 4 * (tid - qr * QPR);
    const bool qrow = qr >= R && qr < R + BY, qcol = ql >= R && ql < R + BX;
    const bool qhas = tid < NQ && (qrow || qcol);
    const int qkeep = qhas && qr < ylim ? min(max(xlim - ql, 0), 4) : 0;
    const long qoff = (long)(y0 - R + qr) * ldx + (x0 - R + ql);
    float4 v1, v2, vb, vm1, vm2, vm3, vm4, vx4, vy4;
    const TileQuads<W, BY, NT> qx(tid, ldx, xlim, ylim_c);
    const TileQuads<BX, H, NT> qy(tid, ldx, xlim_c, ylim);
    const long ox = (long)y0 * ldx + (x0 - R), oy = (long)(y0 - R) * ldx + x0;
    float4 rx1[qx.QT], rx2[qx.QT], ry1[qy.QT], ry2[qy.QT];
    auto fetch = [&](int o) {
        const long oz = (long)o * sz;
        if (qkeep > 0) {
            const long g = oz + qoff;
            v1 = *reinterpret_cast<const float4*>(f1 + g);
            v2 = *reinterpret_cast<const float4*>(f2 + g);
            vb = *reinterpret_cast<const float4*>(b + g);
            vm1 = *reinterpret_cast<const float4*>(m1 + g);
            vm2 = *reinterpret_cast<const float4*>(m2 + g);
            vm3 = *reinterpret_cast<const float4*>(m3 + g);
            vm4 = *reinterpret_cast<const float4*>(m4 + g);
            vx4 = *reinterpret_cast<const float4*>(vrx + g);
            vy4 = *reinterpret_cast<const float4*>(vry + g);
        }
        qx.fetch(rx1, fd1x + oz + ox);
        qx.fetch(rx2, fd2x + oz + ox);
        qy.fetch(ry1, fd1y + oz + oy);
        qy.fetch(ry2, fd2y + oz + oy);
    };
    auto commit = [&](int bf) {
        qx.commit(&fx1[bf][0][0], rx1);
        qx.commit(&fx2[bf][0][0], rx2);
        qy.commit(&fy1[bf][0][0], ry1);
        qy.commit(&fy2[bf][0][0], ry2);
        if (!qhas)
            return;
        const float* A1 = &v1.x;
        const float* A2 = &v2.x;
        const float* BB = &vb.x;
        const float* M1 = &vm1.x;
        const float* M2 = &vm2.x;
        const float* M3 = &vm3.x;
        const float* M4 = &vm4.x;
        const float* VX = &vx4.x;
        const float* VY = &vy4.x;
        float s1x[4], s2x[4], s1y[4], s2y[4], bv[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            s1x[e] = s2x[e] = s1y[e] = s2y[e] = bv[e] = 0.0f;
            if (e < qkeep) {
                const AdjTemps t = adj_temps(A1[e], A2[e], M1[e], M2[e], M3[e], M4[e]);
                const float rho = 1.0f / BB[e];
                const float vx = VX[e], vy = VY[e];
                s1x[e] = rho * (t.lap1 + vx * vx * t.hz1);
                s2x[e] = rho * (t.lap2 + vx * vx * t.hz2);
                s1y[e] = rho * (t.lap1 + vy * vy * t.hz1);
                s2y[e] = rho * (t.lap2 + vy * vy * t.hz2);
                bv[e] = BB[e];
            }
        }
        auto st = [](float* dst, const float* v) {
            *reinterpret_cast<float4*>(dst) = make_float4(v[0], v[1], v[2], v[3]);
        };
        if (qrow) {
            st(&sx1[bf][qr - R][ql], s1x);
            st(&sx2[bf][qr - R][ql], s2x);
        }
        if (qcol) {
            st(&sy1[bf][qr][ql - R], s1y);
            st(&sy2[bf][qr][ql - R], s2y);
        }
        if (qrow && qcol)
            st(&bt[bf][qr - R][ql - R], bv);
    };
    // S_z of this column at plane q (7 words) and Dfwd_z (2 words), loaded one plane before use
    struct Col
    {
        float a1, a2, mm1, mm2, mm3, mm4, bb, vz, d1, d2;
    };
    auto col_load = [&](int q) {
        const long g = (long)q * sz + col;
        return Col{ f1[g], f2[g], m1[g], m2[g], m3[g], m4[g], b[g], vrz[g], fd1z[g], fd2z[g] };
    };
    auto col_sz = [&](const Col& c, float& s1, float& s2) {
        const AdjTemps t = adj_temps(c.a1, c.a2, c.mm1, c.mm2, c.mm3, c.mm4);
        const float rho = 1.0f / c.bb;
        s1 = rho * (t.lap1 + c.vz * c.vz * t.hz1);
        s2 = rho * (t.lap2 + c.vz * c.vz * t.hz2);
    };

    float szq1[9], szq2[9]; // S_z(o - 4 + k), k = 0..8
    float fzq1[8], fzq2[8]; // Dfwd_z(o - 3 + k), k = 0..7
#pragma unroll
    for (int k = 0; k < 9; ++k)
        szq1[k] = szq2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        fzq1[k] = fzq2[k] = 0.0f;
    Col nc = {};
    GT g0 = 0;
    if (active) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const long g = (long)(z0 - 4 + k) * sz + col;
            const AdjTemps t = adj_temps(f1[g], f2[g], m1[g], m2[g], m3[g], m4[g]);
            const float rho = 1.0f / b[g];
            const float vz = vrz[g];
            szq1[k] = rho * (t.lap1 + vz * vz * t.hz1);
            szq2[k] = rho * (t.lap2 + vz * vz * t.hz2);
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            fzq1[k] = fd1z[(long)(z0 - 3 + k) * sz + col];
            fzq2[k] = fd2z[(long)(z0 - 3 + k) * sz + col];
        }
        nc = col_load(z0 + 4);
        g0 = grad[((long)(z0 - izbeg) * gldy + (iy - iybeg)) * gldx + (ix - ixbeg)];
    }
    fetch(z0);
    commit(0);
    if (z0 + 1 < z1)
        fetch(z0 + 1);
    __syncthreads();
    for (int o = z0; o < z1; ++o) {
        const int bf = (o - z0) & 1;
        const GT gold = g0;
        if (active) {
            col_sz(nc, szq1[8], szq2[8]);
            fzq1[7] = nc.d1;
            fzq2[7] = nc.d2;
            if (o + 1 < z1) {
                nc = col_load(o + 5);
                g0 = grad[((long)(o + 1 - izbeg) * gldy + (iy - iybeg)) * gldx + (ix - ixbeg)];
            }
        }
        if (o + 1 < z1)
            commit(bf ^ 1);
        if (o + 2 < z1)
            fetch(o + 2);
        if (active) {
            const float(*SX1)[W] = sx1[bf];
            const float(*SX2)[W] = sx2[bf];
            const float(*SY1)[BX] = sy1[bf];
            const float(*SY2)[BX] = sy2[bf];
            const float(*FX1)[W] = fx1[bf];
            const float(*FX2)[W] = fx2[bf];
            const float(*FY1)[BX] = fy1[bf];
            const float(*FY2)[BX] = fy2[bf];
            float dpx1 = 0, dpx2 = 0, dpy1 = 0, dpy2 = 0, dpz1 = 0, dpz2 = 0; // D+ Dfwd
            float dmx1 = 0, dmx2 = 0, dmy1 = 0, dmy2 = 0, dmz1 = 0, dmz2 = 0; // D- S
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                dpx1 += c * (FX1[ty][tx + R + n] - FX1[ty][tx + R - n + 1]);
                dpx2 += c * (FX2[ty][tx + R + n] - FX2[ty][tx + R - n + 1]);
                dpy1 += c * (FY1[ty + R + n][tx] - FY1[ty + R - n + 1][tx]);
                dpy2 += c * (FY2[ty + R + n][tx] - FY2[ty + R - n + 1][tx]);
                dpz1 += c * (fzq1[3 + n] - fzq1[4 - n]);
                dpz2 += c * (fzq2[3 + n] - fzq2[4 - n]);
                dmx1 += c * (SX1[ty][tx + R + n - 1] - SX1[ty][tx + R - n]);
                dmx2 += c * (SX2[ty][tx + R + n - 1] - SX2[ty][tx + R - n]);
                dmy1 += c * (SY1[ty + R + n - 1][tx] - SY1[ty + R - n][tx]);
                dmy2 += c * (SY2[ty + R + n - 1][tx] - SY2[ty + R - n][tx]);
                dmz1 += c * (szq1[4 + n - 1] - szq1[4 - n]);
                dmz2 += c * (szq2[4 + n - 1] - szq2[4 - n]);
            }
            const float rhs = SX1[ty][tx + R] * (dpx1 * hx) + SX2[ty][tx + R] * (dpx2 * hx) +
                              SY1[ty + R][tx] * (dpy1 * hy) + SY2[ty + R][tx] * (dpy2 * hy) +
                              szq1[4] * (dpz1 * hz) + szq2[4] * (dpz2 * hz) + (dmx1 * hx) * FX1[ty][tx + R] +
                              (dmx2 * hx) * FX2[ty][tx + R] + (dmy1 * hy) * FY1[ty + R][tx] +
                              (dmy2 * hy) * FY2[ty + R][tx] + (dmz1 * hz) * fzq1[3] + (dmz2 * hz) * fzq2[3];
            const long gi = ((long)(o - izbeg) * gldy + (iy - iybeg)) * gldx + (ix - ixbeg);
            grad[gi] = gold + (GT)(-bt[bf][ty][tx] * rhs);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            szq1[k] = szq1[k + 1];
            szq2[k] = szq2[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            fzq1[k] = fzq1[k + 1];
            fzq2[k] = fzq2[k + 1];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Pointwise kernels
// ------------------------------------------------------------------------------------------------------------------

// a4: f <- A f + B dt g over the interior (kernels/kernel_base.cu:82-104)
__global__ __launch_bounds__(256) void k_loop2(float* __restrict__ f1, float* __restrict__ f2,
                                               const float* __restrict__ g1, const float* __restrict__ g2,
                                               const ab2* __restrict__ abx, const ab2* __restrict__ aby,
                                               const ab2* __restrict__ abz, int ixbeg, int ixend, int iybeg,
                                               int izbeg, long ldx, long ldy, float dt)
{
    const int ix = ixbeg + blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = iybeg + blockIdx.y;
    const int iz = izbeg + blockIdx.z;
    if (ix > ixend)
        return;
    const float ayz = fminf(aby[iy].a, abz[iz].a), byz = fminf(aby[iy].b, abz[iz].b);
    const float a = fminf(abx[ix].a, ayz), bb = fminf(abx[ix].b, byz);
    const long i = ((long)iz * ldy + iy) * ldx + ix;
    const float r1 = a * f1[i] + bb * g1[i] * dt;
    const float r2 = a * f2[i] + bb * g2[i] * dt;
    f1[i] = r1;
    f2[i] = r2;
}

// a5: sigma mirror (kernels/kernel_base.cu:142-162)
__global__ __launch_bounds__(256) void k_loop4(float* __restrict__ f1, float* __restrict__ f2, long plane, int izbeg)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (p >= plane)
        return;
    const long i = (long)iz * plane + p;
    if (iz == izbeg) {
        f1[i] = 0.0f;
        f2[i] = 0.0f;
    } else {
        const long m = (long)(2 * izbeg - iz) * plane + p;
        const float r1 = -f1[m], r2 = -f2[m];
        f1[i] = r1;
        f2[i] = r2;
    }
}

// a5: derivative mirror (kernels/kernel_type2_small.cu:5-40)
__global__ __launch_bounds__(256) void k_loop4_derivatives(float* __restrict__ d1x, float* __restrict__ d1y,
                                                           float* __restrict__ d1z, float* __restrict__ d2x,
                                                           float* __restrict__ d2y, float* __restrict__ d2z,
                                                           long plane, int izbeg)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int iz = blockIdx.y;
    if (p >= plane)
        return;
    const long i = (long)iz * plane + p;
    if (iz == izbeg) {
        d1x[i] = 0.0f;
        d1y[i] = 0.0f;
        d2x[i] = 0.0f;
        d2y[i] = 0.0f;
    } else {
        const long m = (long)(2 * izbeg - iz) * plane + p;
        d1x[i] = -d1x[m];
        d1y[i] = -d1y[m];
        d2x[i] = -d2x[m];
        d2y[i] = -d2y[m];
    }
    const long mz = (long)(2 * izbeg + 1 - iz) * plane + p;
    d1z[i] = d1z[mz];
    d2z[i] = d2z[mz];
}

// a8/a13: g(no halo) += adj3*snap3 + adj4*snap4 (kernels/kernel_type2_small.cu:42-59)
__global__ __launch_bounds__(256) void k_vp_grad(const float* __restrict__ s3, const float* __restrict__ s4,
                                                 float* __restrict__ g, const float* __restrict__ a3,
                                                 const float* __restrict__ a4, int ixbeg, int ixend, int iybeg,
                                                 int izbeg, long sldx, long sldy, long ldx, long ldy)
{
    const int ix = ixbeg + blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = iybeg + blockIdx.y;
    const int iz = izbeg + blockIdx.z;
    if (ix > ixend)
        return;
    const long s = ((long)(iz - izbeg) * sldy + (iy - iybeg)) * sldx + (ix - ixbeg);
    const long i = ((long)iz * ldy + iy) * ldx + ix;
    g[s] += a3[i] * s3[s] + a4[i] * s4[s];
}

// finalize: g *= -dt_xcorr*2/vp, vp = sqrt(model3) (kernels/kernel_type2_small.cu:61-76)
__global__ __launch_bounds__(256) void k_vp_grad_finalize(const float* __restrict__ model3, float* __restrict__ g,
                                                          int ixbeg, int ixend, int iybeg, int izbeg, long sldx,
                                                          long sldy, long ldx, long ldy, float dt_xcorr)
{
    const int ix = ixbeg + blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = iybeg + blockIdx.y;
    const int iz = izbeg + blockIdx.z;
    if (ix > ixend)
        return;
    const long s = ((long)(iz - izbeg) * sldy + (iy - iybeg)) * sldx + (ix - ixbeg);
    const float vp = sqrtf(model3[((long)iz * ldy + iy) * ldx + ix]);
    g[s] = (float)((double)g[s] * (-(double)dt_xcorr * 2.0 / (double)vp));
}

// a14: model build from per-point physical parameters (simulator2/simulator2SetUp.cu:10-29)
__global__ __launch_bounds__(256) void k_init_cijs(long n, const float* __restrict__ vp,
                                                   const float* __restrict__ eps, const float* __restrict__ delta,
                                                   const float* __restrict__ vs, const float* __restrict__ theta,
                                                   const float* __restrict__ phi, const float* __restrict__ rho,
                                                   float* __restrict__ m1, float* __restrict__ m2,
                                                   float* __restrict__ m3, float* __restrict__ m4,
                                                   float* __restrict__ m5, float* __restrict__ m6,
                                                   float* __restrict__ m7, float* __restrict__ bb)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float v = vp[i], s = vs[i];
    m1[i] = v * v * (1 + 2 * eps[i]);
    m2[i] = sqrtf((v * v - s * s) * (v * v * (1 + 2 * delta[i]) - s * s));
    m3[i] = v * v;
    m4[i] = s * s;
    const float st = sinf(theta[i]), ct = cosf(theta[i]);
    m5[i] = st * cosf(phi[i]);
    m6[i] = st * sinf(phi[i]);
    m7[i] = ct;
    bb[i] = 1.0f / rho[i];
}

// pressure P = -(2 f1 + f2)/3 (simulator2/simulator2SetUp.cu:31-42)
__global__ __launch_bounds__(256) void k_pressure(long n, const float* __restrict__ f1, const float* __restrict__ f2,
                                                  float* __restrict__ p)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    p[i] = -(2 * f1[i] + f2[i]) / 3;
}

} // namespace tti

// ======================================================================================================================
// C ABI (declared in include/tti_hip.h)
// ======================================================================================================================
using namespace tti;

namespace {

} // namespace

extern "C" {

int
launch_fwd_kernel2_drv(const float* field1, const float* field2, const float* Model11, float* dfield1dx,
                       float* dfield1dy, float* dfield1dz, float* dfield2dx, float* dfield2dy, float* dfield2dz,
                       int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, float invdx, float invdy,
                       float invdz, int ldimx, int ldimy, int order, void* stream)
{
    if (order == 16 && field1 && field2 && Model11 && dfield1dx && dfield1dy && dfield1dz && dfield2dx && dfield2dy &&
        dfield2dz)
        return order_fwd_drv(order, field1, field2, Model11, dfield1dx, dfield1dy, dfield1dz, dfield2dx, dfield2dy,
                             dfield2dz, ixbeg, ixend, iybeg, iyend, izbeg, izend, invdx, invdy, invdz, ldimx, ldimy,
                             stream);
    if (order != 8)
        return order == 16 ? TTI_EINVAL : TTI_EORDER;
    if (!field1 || !field2 || !Model11 || !dfield1dx || !dfield1dy || !dfield1dz || !dfield2dx || !dfield2dy ||
        !dfield2dz || !ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy))
        return TTI_EINVAL;
    const int bz = stencil_bz(ixbeg, ixend, iybeg, iyend, izbeg, izend, BYL);
    auto kern = quads_fit(ixbeg, ldimx, { field1, field2 }) ? k_fwd_drv_q<BYL> : k_fwd_drv<BYL>;
    kern<<<stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz, BYL), dim3(BX, BYL), 0, (hipStream_t)stream>>>(
        field1, field2, Model11, dfield1dx, dfield1dy, dfield1dz, dfield2dx, dfield2dy, dfield2dz, ixbeg, ixend, iybeg,
        iyend, izbeg, izend, invdx, invdy, invdz, (long)ldimx, (long)ldimy, bz);
    return launch_status();
}

int
launch_update_rho_fwd_main_loop_2_1(float* field3, float* field4, const float* field1_dx, const float* field2_dx,
                                    const float* field1_dy, const float* field2_dy, const float* field1_dz,
                                    const float* field2_dz, const float* field1, const float* field2,
                                    const float* irho, const float* model1, const float* model2, const float* model3,
                                    const float* model4, const float* rx, const float* ry, const float* rz,
                                    float* pml_field1_xx, float* pml_field2_xx, float* pml_field1_yy,
                                    float* pml_field2_yy, float* pml_field1_zz, float* pml_field2_zz,
                                    const void* pml_ab_x, const void* pml_ab_y, const void* pml_ab_z,
                                    const void* sponge_ab_xx, const void* sponge_ab_yy, const void* sponge_ab_zz,
                                    int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx,
                                    int ldimy, float dt, double dx, double dy, double dz, int sponge_active,
                                    int order, int pml_x, int pml_y, int pml_z, void* stream)
{
    (void)pml_field1_xx, (void)pml_field2_xx, (void)pml_field1_yy, (void)pml_field2_yy, (void)pml_field1_zz,
        (void)pml_field2_zz, (void)pml_ab_x, (void)pml_ab_y, (void)pml_ab_z, (void)sponge_ab_xx, (void)sponge_ab_yy,
        (void)sponge_ab_zz;
    if (order != 8 && order != 16)
        return TTI_EORDER;
    // The reference instantiates only the interior, no-sponge form of this kernel (kernels/kernel_type2_fwd.cu:641,
    // :657); PML and sponge variants of the forward update are outside the path.
    if (sponge_active || pml_x || pml_y || pml_z)
        return TTI_EINVAL;
    if (!field3 || !field4 || !field1_dx || !field2_dx || !field1_dy || !field2_dy || !field1_dz || !field2_dz ||
        !field1 || !field2 || !irho || !model1 || !model2 || !model3 || !model4 || !rx || !ry || !rz)
        return TTI_EINVAL;
    if (order == 16)
        return order_fwd_update(order, field3, field4, field1_dx, field2_dx, field1_dy, field2_dy, field1_dz,
                                field2_dz, field1, field2, irho, model1, model2, model3, model4, rx, ry, rz, ixbeg,
                                ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, dt, dx, dy, dz, stream);
    if (!ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy))
        return TTI_EINVAL;
    const float invdx = (float)(1.0 / dx), invdy = (float)(1.0 / dy), invdz = (float)(1.0 / dz);
    const float invdxdy = (float)(1.0 / (4 * dx * dy)), invdydz = (float)(1.0 / (4 * dy * dz)),
                invdzdx = (float)(1.0 / (4 * dx * dz));
    const int bz = stencil_bz(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    auto kern = quads_fit(ixbeg, ldimx, { field1, field2, field1_dx, field2_dx, field1_dy, field2_dy })
                    ? k_fwd_update_q<BYS>
                    : k_fwd_update<BYS>;
    kern<<<stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz), dim3(BX, BYS), 0, (hipStream_t)stream>>>(
        field3, field4, field1_dx, field2_dx, field1_dy, field2_dy, field1_dz, field2_dz, field1, field2, irho, model1,
        model2, model3, model4, rx, ry, rz, ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy, dt,
        invdx, invdy, invdz, invdxdy, invdydz, invdzdx, bz);
    return launch_status();
}

int
launch_adj_kernel2_drv(float* dfield1dx, float* dfield2dx, float* dfield1dy, float* dfield2dy, float* dfield1dz,
                       float* dfield2dz, float* field3_rhs, float* field4_rhs, const float* field1,
                       const float* field2, const float* Model11, const float* model1, const float* model2,
                       const float* model3, const float* model4, const float* model5, const float* model6,
                       const float* model7, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                       float invdxx, float invdyy, float invdzz, float invdxy, float invdyz, float invdzx, int ldimx,
                       int ldimy, int order, float dt, float invdx, float invdy, float invdz, void* stream)
{
    (void)invdxx, (void)invdyy, (void)invdzz, (void)dt;
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!dfield1dx || !dfield2dx || !dfield1dy || !dfield2dy || !dfield1dz || !dfield2dz || !field3_rhs ||
        !field4_rhs || !field1 || !field2 || !Model11 || !model1 || !model2 || !model3 || !model4 || !model5 ||
        !model6 || !model7)
        return TTI_EINVAL;
    if (order == 16)
        return order_adj_drv(order, dfield1dx, dfield2dx, dfield1dy, dfield2dy, dfield1dz, dfield2dz, field3_rhs,
                             field4_rhs, field1, field2, Model11, model1, model2, model3, model4, model5, model6,
                             model7, ixbeg, ixend, iybeg, iyend, izbeg, izend, invdxy, invdyz, invdzx, ldimx, ldimy,
                             invdx, invdy, invdz, stream);
    if (!ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy))
        return TTI_EINVAL;
    const int bz = stencil_bz(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    const bool q = quads_fit(ixbeg, ldimx,
                             { field1, field2, Model11, model1, model2, model3, model4, model5, model6, model7 });
    auto kern = q ? k_adj_drv_q<BYS> : k_adj_drv<BYS>;
    kern<<<stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz), dim3(BX, BYS), 0, (hipStream_t)stream>>>(dfield1dx, dfield2dx, dfield1dy, dfield2dy, dfield1dz, dfield2dz,
                       field3_rhs, field4_rhs, field1, field2, Model11, model1, model2, model3, model4, model5,
                       model6, model7, ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy, invdx,
                       invdy, invdz, invdxy, invdyz, invdzx, bz);
    return launch_status();
}

int
launch_update_adj_kernel_main_3(float* field3, float* field4, const float* field1_dx, const float* field2_dx,
                                const float* field1_dy, const float* field2_dy, const float* field1_dz,
                                const float* field2_dz, const float* field3_rhs, const float* field4_rhs,
                                const void* ab_xx, const void* ab_yy, const void* ab_zz, int ixbeg, int ixend,
                                int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy, float dt, double dx,
                                double dy, double dz, int pmlx, int pmly, int pmlz, int order, int sponge_active,
                                void* stream)
{
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (pmlx || pmly || pmlz)
        return TTI_EINVAL; // PML is outside the path (the reference drivers zero every PML pointer)
    if (!field3 || !field4 || !field1_dx || !field2_dx || !field1_dy || !field2_dy || !field1_dz || !field2_dz ||
        !field3_rhs || !field4_rhs)
        return TTI_EINVAL;
    if (sponge_active && (!ab_xx || !ab_yy || !ab_zz))
        return TTI_EINVAL;
    if (order == 16)
        return order_adj_update(order, field3, field4, field1_dx, field2_dx, field1_dy, field2_dy, field1_dz,
                                field2_dz, field3_rhs, field4_rhs, ab_xx, ab_yy, ab_zz, ixbeg, ixend, iybeg, iyend,
                                izbeg, izend, ldimx, ldimy, dt, dx, dy, dz, sponge_active, stream);
    if (!ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy))
        return TTI_EINVAL;
    const float invdx = (float)(1.0 / dx), invdy = (float)(1.0 / dy), invdz = (float)(1.0 / dz);
    const int bz = BZ_DEFAULT; // plain dispatch order, 64-plane chunks: the fastest measured for this kernel
    const dim3 grid = stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz, BYL);
    const bool q = quads_fit(ixbeg, ldimx, { field1_dx, field2_dx, field1_dy, field2_dy });
    auto kern = sponge_active ? (q ? k_adj_update_q<BYL, true> : k_adj_update<BYL, true>)
                              : (q ? k_adj_update_q<BYL, false> : k_adj_update<BYL, false>);
    const ab2* ab[3] = { nullptr, nullptr, nullptr };
    if (sponge_active) {
        ab[0] = (const ab2*)ab_xx;
        ab[1] = (const ab2*)ab_yy;
        ab[2] = (const ab2*)ab_zz;
    }
    kern<<<grid, dim3(BX, BYL), 0, (hipStream_t)stream>>>(field3, field4, field1_dx, field2_dx, field1_dy, field2_dy,
                                                          field1_dz, field2_dz, field3_rhs, field4_rhs, ab[0], ab[1],
                                                          ab[2], ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx,
                                                          (long)ldimy, dt, invdx, invdy, invdz, bz);
    return launch_status();
}

int
launch_update_rho_gradient_rho(float* grad, const float* fwd_field1_dx, const float* fwd_field2_dx,
                               const float* fwd_field1_dy, const float* fwd_field2_dy, const float* fwd_field1_dz,
                               const float* fwd_field2_dz, const float* field1, const float* field2,
                               const float* irho, const float* model1, const float* model2, const float* model3,
                               const float* model4, const float* rx, const float* ry, const float* rz, int ixbeg,
                               int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy,
                               int grad_ldimx, int grad_ldimy, double dx, double dy, double dz, int order,
                               int simple_gradient, void* stream)
{
    (void)simple_gradient; // both reference kernels compute the same quantity; one gfx950 kernel serves both
    if (order != 8 && order != 12 && order != 16)
        return TTI_EORDER;
    if (ixend < ixbeg || iyend < iybeg || izend < izbeg)
        return TTI_OK; // empty range: nothing to do (kernels/kernel_type2_grad.cu:625-626)
    if (!grad || !fwd_field1_dx || !fwd_field2_dx || !fwd_field1_dy || !fwd_field2_dy || !fwd_field1_dz ||
        !fwd_field2_dz || !field1 || !field2 || !irho || !model1 || !model2 || !model3 || !model4 || !rx || !ry ||
        !rz || grad_ldimx < ixend - ixbeg + 1 || grad_ldimy < iyend - iybeg + 1)
        return TTI_EINVAL;
    if (order != 8)
        return order_rho_grad(order, grad, fwd_field1_dx, fwd_field2_dx, fwd_field1_dy, fwd_field2_dy, fwd_field1_dz,
                              fwd_field2_dz, field1, field2, irho, model1, model2, model3, model4, rx, ry, rz, ixbeg,
                              ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, grad_ldimx, grad_ldimy, dx, dy, dz,
                              stream);
    if (!ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy))
        return TTI_EINVAL;
    const float hx = (float)(1.0 / dx), hy = (float)(1.0 / dy), hz = (float)(1.0 / dz);
    const int bz = stencil_bz(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    const bool q = quads_fit(ixbeg, ldimx, { fwd_field1_dx, fwd_field2_dx, fwd_field1_dy, fwd_field2_dy, field1, field2,
                                            irho, model1, model2, model3, model4, rx, ry });
    auto kern = q ? k_rho_grad_q<BYS> : k_rho_grad<BYS>;
    kern<<<stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz), dim3(BX, BYS), 0, (hipStream_t)stream>>>(
        grad, fwd_field1_dx, fwd_field2_dx, fwd_field1_dy, fwd_field2_dy, fwd_field1_dz, fwd_field2_dz, field1, field2,
        irho, model1, model2, model3, model4, rx, ry, rz, ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx,
        (long)ldimy, (long)grad_ldimx, (long)grad_ldimy, hx, hy, hz, bz);
    return launch_status();
}

} // extern "C"

namespace tti {

int
rho_gradient_f64(double* grad, const float* fwd_field1_dx, const float* fwd_field2_dx, const float* fwd_field1_dy,
                 const float* fwd_field2_dy, const float* fwd_field1_dz, const float* fwd_field2_dz,
                 const float* field1, const float* field2, const float* irho, const float* model1,
                 const float* model2, const float* model3, const float* model4, const float* rx, const float* ry,
                 const float* rz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx,
                 int ldimy, int grad_ldimx, int grad_ldimy, double dx, double dy, double dz, void* stream)
{
    if (!grad || !fwd_field1_dx || !fwd_field2_dx || !fwd_field1_dy || !fwd_field2_dy || !fwd_field1_dz ||
        !fwd_field2_dz || !field1 || !field2 || !irho || !model1 || !model2 || !model3 || !model4 || !rx || !ry ||
        !rz || !ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        grad_ldimx < ixend - ixbeg + 1 || grad_ldimy < iyend - iybeg + 1)
        return TTI_EINVAL;
    const float hx = (float)(1.0 / dx), hy = (float)(1.0 / dy), hz = (float)(1.0 / dz);
    const int bz = stencil_bz(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    const bool q = quads_fit(ixbeg, ldimx, { fwd_field1_dx, fwd_field2_dx, fwd_field1_dy, fwd_field2_dy, field1, field2,
                                            irho, model1, model2, model3, model4, rx, ry });
    auto kern = q ? k_rho_grad_q<BYS, double> : k_rho_grad<BYS, double>;
    kern<<<stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz), dim3(BX, BYS), 0, (hipStream_t)stream>>>(
        grad, fwd_field1_dx, fwd_field2_dx, fwd_field1_dy, fwd_field2_dy, fwd_field1_dz, fwd_field2_dz, field1, field2,
        irho, model1, model2, model3, model4, rx, ry, rz, ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx,
        (long)ldimy, (long)grad_ldimx, (long)grad_ldimy, hx, hy, hz, bz);
    return launch_status();
}

} // namespace tti

extern "C" {

int
kernel_loop2_inner(float* field1, float* field2, const float* field3, const float* field4, const void* ABx,
                   const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                   int ldimx, int ldimy, float dt, void* stream)
{
    if (!field1 || !field2 || !field3 || !field4 || !ABx || !ABy || !ABz || ixend < ixbeg || iyend < iybeg ||
        izend < izbeg || ixbeg < 0 || iybeg < 0 || izbeg < 0 || ixend >= ldimx || iyend >= ldimy)
        return TTI_EINVAL;
    const int nx = ixend - ixbeg + 1;
    dim3 grid((nx + 255) / 256, iyend - iybeg + 1, izend - izbeg + 1);
    k_loop2<<<grid, dim3(256), 0, (hipStream_t)stream>>>(field1, field2, field3, field4,
                       (const ab2*)ABx, (const ab2*)ABy, (const ab2*)ABz, ixbeg, ixend, iybeg, izbeg, (long)ldimx,
                       (long)ldimy, dt);
    return launch_status();
}

int
kernel_loop4_inner(float* field1, float* field2, int ldimx, int ldimy, int izbeg, void* stream)
{
    if (!field1 || !field2 || ldimx <= 0 || ldimy <= 0 || izbeg < 0)
        return TTI_EINVAL;
    const long plane = (long)ldimx * ldimy;
    dim3 grid((unsigned)((plane + 255) / 256), izbeg + 1);
    k_loop4<<<grid, dim3(256), 0, (hipStream_t)stream>>>(field1, field2, plane, izbeg);
    return launch_status();
}

int
kernel_loop4_derivatives(float* dfield1dx, float* dfield1dy, float* dfield1dz, float* dfield2dx, float* dfield2dy,
                         float* dfield2dz, int ldimx, int ldimy, int izbeg, void* stream)
{
    if (!dfield1dx || !dfield1dy || !dfield1dz || !dfield2dx || !dfield2dy || !dfield2dz || ldimx <= 0 ||
        ldimy <= 0 || izbeg < 0)
        return TTI_EINVAL;
    const long plane = (long)ldimx * ldimy;
    dim3 grid((unsigned)((plane + 255) / 256), izbeg + 1);
    k_loop4_derivatives<<<grid, dim3(256), 0, (hipStream_t)stream>>>(dfield1dx, dfield1dy,
                       dfield1dz, dfield2dx, dfield2dy, dfield2dz, plane, izbeg);
    return launch_status();
}

int
update_Vp_gradient(const float* snap_field3, const float* snap_field4, float* grad_Vp, const float* adj_field3,
                   const float* adj_field4, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                   int snap_ldimx, int snap_ldimy, int ldimx, int ldimy, void* stream)
{
    if (!snap_field3 || !snap_field4 || !grad_Vp || !adj_field3 || !adj_field4 || ixend < ixbeg || iyend < iybeg ||
        izend < izbeg || snap_ldimx < ixend - ixbeg + 1 || snap_ldimy < iyend - iybeg + 1 || ixend >= ldimx ||
        iyend >= ldimy)
        return TTI_EINVAL;
    dim3 grid((ixend - ixbeg + 256) / 256, iyend - iybeg + 1, izend - izbeg + 1);
    k_vp_grad<<<grid, dim3(256), 0, (hipStream_t)stream>>>(snap_field3, snap_field4, grad_Vp,
                       adj_field3, adj_field4, ixbeg, ixend, iybeg, izbeg, (long)snap_ldimx, (long)snap_ldimy,
                       (long)ldimx, (long)ldimy);
    return launch_status();
}

int
finalize_Vp_gradient(const float* model3, float* grad_Vp, int ixbeg, int ixend, int iybeg, int iyend, int izbeg,
                     int izend, int grad_ldimx, int grad_ldimy, int ldimx, int ldimy, float dt_xcorr, void* stream)
{
    if (!model3 || !grad_Vp || ixend < ixbeg || iyend < iybeg || izend < izbeg ||
        grad_ldimx < ixend - ixbeg + 1 || grad_ldimy < iyend - iybeg + 1 || ixend >= ldimx || iyend >= ldimy)
        return TTI_EINVAL;
    dim3 grid((ixend - ixbeg + 256) / 256, iyend - iybeg + 1, izend - izbeg + 1);
    k_vp_grad_finalize<<<grid, dim3(256), 0, (hipStream_t)stream>>>(model3, grad_Vp, ixbeg, ixend,
                       iybeg, izbeg, (long)grad_ldimx, (long)grad_ldimy, (long)ldimx, (long)ldimy, dt_xcorr);
    return launch_status();
}

int
init_cijs(long n, const float* vp, const float* eps, const float* delta, const float* vs, const float* theta,
          const float* phi, const float* rho, float* model1, float* model2, float* model3, float* model4,
          float* model5, float* model6, float* model7, float* Model11, void* stream)
{
    if (n < 0 || (n > 0 && (!vp || !eps || !delta || !vs || !theta || !phi || !rho || !model1 || !model2 ||
                            !model3 || !model4 || !model5 || !model6 || !model7 || !Model11)))
        return TTI_EINVAL;
    if (n == 0)
        return TTI_OK;
    k_init_cijs<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream>>>(n, vp, eps,
                       delta, vs, theta, phi, rho, model1, model2, model3, model4, model5, model6, model7, Model11);
    return launch_status();
}

int
compute_pressure(long n, const float* field1, const float* field2, float* pressure, void* stream)
{
    if (n < 0 || (n > 0 && (!field1 || !field2 || !pressure)))
        return TTI_EINVAL;
    if (n == 0)
        return TTI_OK;
    k_pressure<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream>>>(n, field1,
                       field2, pressure);
    return launch_status();
}

} // extern "C"
