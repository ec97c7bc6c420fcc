// Constant-density TTI operator (the "type1" family) for gfx950: forward, adjoint and imaging condition.
//
// Central second derivatives F_aa = (1/da^2)(e0 f + sum_n e_n (f(-n) + f(+n))) replace the staggered D+ b D- pair of
// the variable-density operator; the rotated mixed derivatives and the V/H model combination are the same
// (kernels/kernel_type1_gpu_fwd.cu:9-71, :225-437; kernels/kernel_type1_gpu_adj.cu:9-51, :235-378;
// kernels/helper_kernels_gpu.h:120-418). Same 2.5-D structure as the type2 kernels: 64 x 8 tiles in LDS, a z-march
// with the z stencil in a 9-deep register queue, and plane p = o + 4 absorbed into the x-z / y-z partial sums.
#include "order_kernels.h"
#include "stencil_common.h"

namespace tti {

__device__ __forceinline__ float
central2(float c, float m1, float p1, float m2, float p2, float m3, float p3, float m4, float p4)
{
    return E0 * c + E1 * (m1 + p1) + E2 * (m2 + p2) + E3 * (m3 + p3) + E4 * (m4 + p4);
}

// MODE 0: u += dt*rhs; MODE 1: u = A u + B dt rhs (sponge); MODE 2: out(no halo) = rhs (calc_snap).
template <int BY, int MODE>
__global__ __launch_bounds__(BX* BY) void k_t1_fwd(
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ f1, const float* __restrict__ f2,
    const float* __restrict__ m1, const float* __restrict__ m2, const float* __restrict__ m3,
    const float* __restrict__ m4, const float* __restrict__ m5, const float* __restrict__ m6,
    const float* __restrict__ m7, const ab2* __restrict__ abx, const ab2* __restrict__ aby,
    const ab2* __restrict__ abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx,
    long ldy, long oldx, long oldy, float dt, float invdxx, float invdyy, float invdzz, float i4xy, float i4yz,
    float i4zx, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R;
    __shared__ float s1[H][W], s2[H][W];
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

    float cq1[9], cq2[9];                            // sigma(p - 8 + k)
    float xx1[5], xx2[5], yy1[5], yy2[5], xy1[5], xy2[5]; // F(p - 4 + k)
    float qx1[9], qx2[9], qy1[9], qy2[9];            // partial sums, q[j] <-> output p - 4 + j
#pragma unroll
    for (int k = 0; k < 9; ++k)
        cq1[k] = cq2[k] = qx1[k] = qx2[k] = qy1[k] = qy2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        xx1[k] = xx2[k] = yy1[k] = yy2[k] = xy1[k] = xy2[k] = 0.0f;

    for (int p = z0 - R; p < z1 + R; ++p) {
        const int o = p - R;
        const long org = (long)p * sz + (long)(y0 - R) * ldx + (x0 - R);
        load_tile<BY, W, H>(s1, f1, org, ldx, xlim, ylim);
        load_tile<BY, W, H>(s2, f2, org, ldx, xlim, ylim);
        __syncthreads();
        if (active) {
            const int cy = ty + R, cx = tx + R;
            cq1[8] = s1[cy][cx];
            cq2[8] = s2[cy][cx];
            xx1[4] = invdxx * central2(s1[cy][cx], s1[cy][cx - 1], s1[cy][cx + 1], s1[cy][cx - 2], s1[cy][cx + 2],
                                       s1[cy][cx - 3], s1[cy][cx + 3], s1[cy][cx - 4], s1[cy][cx + 4]);
            xx2[4] = invdxx * central2(s2[cy][cx], s2[cy][cx - 1], s2[cy][cx + 1], s2[cy][cx - 2], s2[cy][cx + 2],
                                       s2[cy][cx - 3], s2[cy][cx + 3], s2[cy][cx - 4], s2[cy][cx + 4]);
            yy1[4] = invdyy * central2(s1[cy][cx], s1[cy - 1][cx], s1[cy + 1][cx], s1[cy - 2][cx], s1[cy + 2][cx],
                                       s1[cy - 3][cx], s1[cy + 3][cx], s1[cy - 4][cx], s1[cy + 4][cx]);
            yy2[4] = invdyy * central2(s2[cy][cx], s2[cy - 1][cx], s2[cy + 1][cx], s2[cy - 2][cx], s2[cy + 2][cx],
                                       s2[cy - 3][cx], s2[cy + 3][cx], s2[cy - 4][cx], s2[cy + 4][cx]);
            float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float e = ecen(n);
                a1 += e * ((s1[cy + n][cx + n] - s1[cy - n][cx + n]) + (s1[cy - n][cx - n] - s1[cy + n][cx - n]));
                a2 += e * ((s2[cy + n][cx + n] - s2[cy - n][cx + n]) + (s2[cy - n][cx - n] - s2[cy + n][cx - n]));
                const float dx1 = s1[cy][cx + n] - s1[cy][cx - n], dx2 = s2[cy][cx + n] - s2[cy][cx - n];
                const float dy1 = s1[cy + n][cx] - s1[cy - n][cx], dy2 = s2[cy + n][cx] - s2[cy - n][cx];
                qx1[4 - n] += e * dx1;
                qx1[4 + n] -= e * dx1;
                qx2[4 - n] += e * dx2;
                qx2[4 + n] -= e * dx2;
                qy1[4 - n] += e * dy1;
                qy1[4 + n] -= e * dy1;
                qy2[4 - n] += e * dy2;
                qy2[4 + n] -= e * dy2;
            }
            xy1[4] = a1 * i4xy;
            xy2[4] = a2 * i4xy;
            if (o >= z0) {
                const long i = (long)o * sz + col;
                const float fzz1 = invdzz * central2(cq1[4], cq1[3], cq1[5], cq1[2], cq1[6], cq1[1], cq1[7], cq1[0],
                                                     cq1[8]);
                const float fzz2 = invdzz * central2(cq2[4], cq2[3], cq2[5], cq2[2], cq2[6], cq2[1], cq2[7], cq2[0],
                                                     cq2[8]);
                const float vx = m5[i], vy = m6[i], vz = m7[i];
                const float rx2 = vx * vx, ry2 = vy * vy, rz2 = vz * vz;
                const float rxy = 2.0f * vx * vy, ryz = 2.0f * vy * vz, rzx = 2.0f * vz * vx;
                const float f1zx = qx1[0] * i4zx, f2zx = qx2[0] * i4zx, f1yz = qy1[0] * i4yz, f2yz = qy2[0] * i4yz;
                const float V1 = rx2 * xx1[0] + ry2 * yy1[0] + rz2 * fzz1 + rxy * xy1[0] + rzx * f1zx + ryz * f1yz;
                const float H1 = xx1[0] + yy1[0] + fzz1 - V1;
                const float V2 = rx2 * xx2[0] + ry2 * yy2[0] + rz2 * fzz2 + rxy * xy2[0] + rzx * f2zx + ryz * f2yz;
                const float H2 = xx2[0] + yy2[0] + fzz2 - V2;
                const float mm1 = m1[i], mm2 = m2[i], mm3 = m3[i], mm4 = m4[i];
                const float r3 = mm1 * H1 + mm4 * V1 + mm2 * V2;
                const float r4 = mm2 * H1 + mm3 * V2 + mm4 * H2;
                if (MODE == 2) {
                    const long oi = ((long)(o - izbeg) * oldy + (iy - iybeg)) * oldx + (ix - ixbeg);
                    u[oi] = r3;
                    w[oi] = r4;
                } else if (MODE == 1) {
                    const float a = fminf(abx[ix].a, fminf(aby[iy].a, abz[o].a));
                    const float bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[o].b));
                    u[i] = a * u[i] + bb * r3 * dt;
                    w[i] = a * w[i] + bb * r4 * dt;
                } else {
                    u[i] += r3 * dt;
                    w[i] += r4 * dt;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            cq1[k] = cq1[k + 1];
            cq2[k] = cq2[k + 1];
            qx1[k] = qx1[k + 1];
            qx2[k] = qx2[k + 1];
            qy1[k] = qy1[k + 1];
            qy2[k] = qy2[k + 1];
        }
        qx1[8] = qx2[8] = qy1[8] = qy2[8] = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            xx1[k] = xx1[k + 1];
            xx2[k] = xx2[k + 1];
            yy1[k] = yy1[k + 1];
            yy2[k] = yy2[k + 1];
            xy1[k] = xy1[k + 1];
            xy2[k] = xy2[k + 1];
        }
    }
}

// Transposed constant-density operator:
//   r = sum_a central2_a(T_a) + sum_{a<b} X_ab(M_ab),  T_{f,a} = lap_f + r_a^2 hz_f,  M_{f,ab} = 2 r_a r_b hz_f
//   u += dt r  or, with the sponge, u = A u + B dt r.
template <int BY, bool SPONGE>
__global__ __launch_bounds__(BX* BY) void k_t1_adj(
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ f1, const float* __restrict__ f2,
    const float* __restrict__ m1, const float* __restrict__ m2, const float* __restrict__ m3,
    const float* __restrict__ m4, const float* __restrict__ m5, const float* __restrict__ m6,
    const float* __restrict__ m7, const ab2* __restrict__ abx, const ab2* __restrict__ aby,
    const ab2* __restrict__ abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx,
    long ldy, float dt, float invdxx, float invdyy, float invdzz, float i4xy, float i4yz, float i4zx, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R;
    __shared__ float tx1[BY][W], tx2[BY][W];   // T_x (x halo)
    __shared__ float ty1[H][BX], ty2[H][BX];   // T_y (y halo)
    __shared__ float mxy1[H][W], mxy2[H][W];   // M_xy
    __shared__ float mzx1[BY][W], mzx2[BY][W]; // M_zx (x halo)
    __shared__ float myz1[H][BX], myz2[H][BX]; // M_yz (y halo)
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

    float cq1[9], cq2[9];                                  // T_z(p - 8 + k)
    float cx1[5], cx2[5], cy1[5], cy2[5], xy1[5], xy2[5]; // central2_x(T_x), central2_y(T_y), X_xy(M_xy) at p-4+k
    float q1[9], q2[9];                                    // X_zx + X_yz partial sums
#pragma unroll
    for (int k = 0; k < 9; ++k)
        cq1[k] = cq2[k] = q1[k] = q2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        cx1[k] = cx2[k] = cy1[k] = cy2[k] = xy1[k] = xy2[k] = 0.0f;

    for (int p = z0 - R; p < z1 + R; ++p) {
        const int o = p - R;
        for (int k = ty * BX + tx; k < W * H; k += BX * BY) {
            const int ly = k / W, lx = k - ly * W;
            float vx1 = 0, vx2 = 0, vy1 = 0, vy2 = 0, vxy1 = 0, vxy2 = 0, vzx1 = 0, vzx2 = 0, vyz1 = 0, vyz2 = 0;
            if (lx < xlim && ly < ylim) {
                const long g = (long)p * sz + (long)(y0 - R + ly) * ldx + (x0 - R + lx);
                const float a = f1[g], c = f2[g], mm1 = m1[g], mm2 = m2[g], mm3 = m3[g], mm4 = m4[g];
                const float lap1 = mm1 * a + mm2 * c, hz1 = (-mm1 + mm4) * a - mm2 * c;
                const float lap2 = mm4 * c, hz2 = mm2 * a + (-mm4 + mm3) * c;
                const float rx = m5[g], ry = m6[g], rz = m7[g];
                vx1 = lap1 + rx * rx * hz1;
                vx2 = lap2 + rx * rx * hz2;
                vy1 = lap1 + ry * ry * hz1;
                vy2 = lap2 + ry * ry * hz2;
                vxy1 = 2.0f * rx * ry * hz1;
                vxy2 = 2.0f * rx * ry * hz2;
                vzx1 = 2.0f * rz * rx * hz1;
                vzx2 = 2.0f * rz * rx * hz2;
                vyz1 = 2.0f * ry * rz * hz1;
                vyz2 = 2.0f * ry * rz * hz2;
            }
            mxy1[ly][lx] = vxy1;
            mxy2[ly][lx] = vxy2;
            if (ly >= R && ly < R + BY) {
                tx1[ly - R][lx] = vx1;
                tx2[ly - R][lx] = vx2;
                mzx1[ly - R][lx] = vzx1;
                mzx2[ly - R][lx] = vzx2;
            }
            if (lx >= R && lx < R + BX) {
                ty1[ly][lx - R] = vy1;
                ty2[ly][lx - R] = vy2;
                myz1[ly][lx - R] = vyz1;
                myz2[ly][lx - R] = vyz2;
            }
        }
        if (active) {
            // own column T_z(p)
            const long g = (long)p * sz + col;
            const float a = f1[g], c = f2[g], mm1 = m1[g], mm2 = m2[g], mm3 = m3[g], mm4 = m4[g], rz = m7[g];
            cq1[8] = mm1 * a + mm2 * c + rz * rz * ((-mm1 + mm4) * a - mm2 * c);
            cq2[8] = mm4 * c + rz * rz * (mm2 * a + (-mm4 + mm3) * c);
        }
        __syncthreads();
        if (active) {
            const int cx = tx + R, cy = ty + R;
            cx1[4] = invdxx * central2(tx1[ty][cx], tx1[ty][cx - 1], tx1[ty][cx + 1], tx1[ty][cx - 2],
                                       tx1[ty][cx + 2], tx1[ty][cx - 3], tx1[ty][cx + 3], tx1[ty][cx - 4],
                                       tx1[ty][cx + 4]);
            cx2[4] = invdxx * central2(tx2[ty][cx], tx2[ty][cx - 1], tx2[ty][cx + 1], tx2[ty][cx - 2],
                                       tx2[ty][cx + 2], tx2[ty][cx - 3], tx2[ty][cx + 3], tx2[ty][cx - 4],
                                       tx2[ty][cx + 4]);
            cy1[4] = invdyy * central2(ty1[cy][tx], ty1[cy - 1][tx], ty1[cy + 1][tx], ty1[cy - 2][tx],
                                       ty1[cy + 2][tx], ty1[cy - 3][tx], ty1[cy + 3][tx], ty1[cy - 4][tx],
                                       ty1[cy + 4][tx]);
            cy2[4] = invdyy * central2(ty2[cy][tx], ty2[cy - 1][tx], ty2[cy + 1][tx], ty2[cy - 2][tx],
                                       ty2[cy + 2][tx], ty2[cy - 3][tx], ty2[cy + 3][tx], ty2[cy - 4][tx],
                                       ty2[cy + 4][tx]);
            float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float e = ecen(n);
                a1 += e * (mxy1[cy + n][cx + n] - mxy1[cy + n][cx - n] + mxy1[cy - n][cx - n] - mxy1[cy - n][cx + n]);
                a2 += e * (mxy2[cy + n][cx + n] - mxy2[cy + n][cx - n] + mxy2[cy - n][cx - n] - mxy2[cy - n][cx + n]);
                const float d1 = i4zx * (mzx1[ty][cx + n] - mzx1[ty][cx - n]) + i4yz * (myz1[cy + n][tx] - myz1[cy - n][tx]);
                const float d2 = i4zx * (mzx2[ty][cx + n] - mzx2[ty][cx - n]) + i4yz * (myz2[cy + n][tx] - myz2[cy - n][tx]);
                q1[4 - n] += e * d1;
                q1[4 + n] -= e * d1;
                q2[4 - n] += e * d2;
                q2[4 + n] -= e * d2;
            }
            xy1[4] = a1 * i4xy;
            xy2[4] = a2 * i4xy;
            if (o >= z0) {
                const long i = (long)o * sz + col;
                const float cz1 = invdzz * central2(cq1[4], cq1[3], cq1[5], cq1[2], cq1[6], cq1[1], cq1[7], cq1[0],
                                                    cq1[8]);
                const float cz2 = invdzz * central2(cq2[4], cq2[3], cq2[5], cq2[2], cq2[6], cq2[1], cq2[7], cq2[0],
                                                    cq2[8]);
                const float r3 = cx1[0] + cy1[0] + cz1 + xy1[0] + q1[0];
                const float r4 = cx2[0] + cy2[0] + cz2 + xy2[0] + q2[0];
                if (SPONGE) {
                    const float a = fminf(abx[ix].a, fminf(aby[iy].a, abz[o].a));
                    const float bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[o].b));
                    u[i] = a * u[i] + bb * r3 * dt;
                    w[i] = a * w[i] + bb * r4 * dt;
                } else {
                    u[i] += r3 * dt;
                    w[i] += r4 * dt;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            cq1[k] = cq1[k + 1];
            cq2[k] = cq2[k + 1];
            q1[k] = q1[k + 1];
            q2[k] = q2[k + 1];
        }
        q1[8] = q2[8] = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cx1[k] = cx1[k + 1];
            cx2[k] = cx2[k + 1];
            cy1[k] = cy1[k + 1];
            cy2[k] = cy2[k + 1];
            xy1[k] = xy1[k + 1];
            xy2[k] = xy2[k + 1];
        }
    }
}

// The forward operator with the plane pipeline of the gfx950 stencils: the sigma tiles of plane p + 1 are loaded as
// float4 quads during plane p - 1 and stored into the other LDS buffer during plane p; the pointwise words of the next
// output plane are loaded one plane ahead; one barrier per plane. Needs quads_fit (the launcher falls back).
template <int BY, int MODE>
__global__ __launch_bounds__(BX* BY) void k_t1_fwd_q(
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ f1, const float* __restrict__ f2,
    const float* __restrict__ m1, const float* __restrict__ m2, const float* __restrict__ m3,
    const float* __restrict__ m4, const float* __restrict__ m5, const float* __restrict__ m6,
    const float* __restrict__ m7, const ab2* __restrict__ abx, const ab2* __restrict__ aby,
    const ab2* __restrict__ abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx,
    long ldy, long oldx, long oldy, float dt, float invdxx, float invdyy, float invdzz, float i4xy, float i4yz,
    float i4zx, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R, NT = BX * BY;
    __shared__ __align__(16) float s1[2][H][W], s2[2][H][W];
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
    const TileQuads<W, H, NT> qt(tid, ldx, xlim, ylim);
    const long org = (long)(y0 - R) * ldx + (x0 - R);
    float4 r1[qt.QT], r2[qt.QT];
    auto fetch = [&](int p) {
        qt.fetch(r1, f1 + (long)p * sz + org);
        qt.fetch(r2, f2 + (long)p * sz + org);
    };
    float axy = 1.0f, bxy = 1.0f;
    if (MODE == 1 && active) {
        axy = fminf(abx[ix].a, aby[iy].a);
        bxy = fminf(abx[ix].b, aby[iy].b);
    }
    float pw[9]; // m1..m7, u, w of the output plane being finalised
    auto load_pw = [&](int o) {
        if (active && o >= z0 && o < z1) {
            const long i = (long)o * sz + col;
            pw[0] = m1[i];
            pw[1] = m2[i];
            pw[2] = m3[i];
            pw[3] = m4[i];
            pw[4] = m5[i];
            pw[5] = m6[i];
            pw[6] = m7[i];
            if (MODE != 2) {
                pw[7] = u[i];
                pw[8] = w[i];
            }
        }
    };

    float cq1[9], cq2[9];                            // sigma(p - 8 + k)
    float xx1[5], xx2[5], yy1[5], yy2[5], xy1[5], xy2[5]; // F(p - 4 + k)
    float qx1[9], qx2[9], qy1[9], qy2[9];            // partial sums, q[j] <-> output p - 4 + j
#pragma unroll
    for (int k = 0; k < 9; ++k)
        cq1[k] = cq2[k] = qx1[k] = qx2[k] = qy1[k] = qy2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        xx1[k] = xx2[k] = yy1[k] = yy2[k] = xy1[k] = xy2[k] = 0.0f;

    load_pw(p0 - R + 1);
    fetch(p0);
    qt.commit(&s1[0][0][0], r1);
    qt.commit(&s2[0][0][0], r2);
    if (p0 + 1 < pend)
        fetch(p0 + 1);
    __syncthreads();
    for (int p = p0; p < pend; ++p) {
        const int o = p - R, bf = (p - p0) & 1;
        if (p + 1 < pend) {
            qt.commit(&s1[bf ^ 1][0][0], r1);
            qt.commit(&s2[bf ^ 1][0][0], r2);
        }
        if (p + 2 < pend)
            fetch(p + 2);
        float un = 0.0f, wn = 0.0f;
        if (active) {
            const float(*S1)[W] = s1[bf];
            const float(*S2)[W] = s2[bf];
            const int cy = ty + R, cx = tx + R;
            cq1[8] = S1[cy][cx];
            cq2[8] = S2[cy][cx];
            xx1[4] = invdxx * central2(S1[cy][cx], S1[cy][cx - 1], S1[cy][cx + 1], S1[cy][cx - 2], S1[cy][cx + 2],
                                       S1[cy][cx - 3], S1[cy][cx + 3], S1[cy][cx - 4], S1[cy][cx + 4]);
            xx2[4] = invdxx * central2(S2[cy][cx], S2[cy][cx - 1], S2[cy][cx + 1], S2[cy][cx - 2], S2[cy][cx + 2],
                                       S2[cy][cx - 3], S2[cy][cx + 3], S2[cy][cx - 4], S2[cy][cx + 4]);
            yy1[4] = invdyy * central2(S1[cy][cx], S1[cy - 1][cx], S1[cy + 1][cx], S1[cy - 2][cx], S1[cy + 2][cx],
                                       S1[cy - 3][cx], S1[cy + 3][cx], S1[cy - 4][cx], S1[cy + 4][cx]);
            yy2[4] = invdyy * central2(S2[cy][cx], S2[cy - 1][cx], S2[cy + 1][cx], S2[cy - 2][cx], S2[cy + 2][cx],
                                       S2[cy - 3][cx], S2[cy + 3][cx], S2[cy - 4][cx], S2[cy + 4][cx]);
            float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float e = ecen(n);
                a1 += e * ((S1[cy + n][cx + n] - S1[cy - n][cx + n]) + (S1[cy - n][cx - n] - S1[cy + n][cx - n]));
                a2 += e * ((S2[cy + n][cx + n] - S2[cy - n][cx + n]) + (S2[cy - n][cx - n] - S2[cy + n][cx - n]));
                const float dx1 = S1[cy][cx + n] - S1[cy][cx - n], dx2 = S2[cy][cx + n] - S2[cy][cx - n];
                const float dy1 = S1[cy + n][cx] - S1[cy - n][cx], dy2 = S2[cy + n][cx] - S2[cy - n][cx];
                qx1[4 - n] += e * dx1;
                qx1[4 + n] -= e * dx1;
                qx2[4 - n] += e * dx2;
                qx2[4 + n] -= e * dx2;
                qy1[4 - n] += e * dy1;
                qy1[4 + n] -= e * dy1;
                qy2[4 - n] += e * dy2;
                qy2[4 + n] -= e * dy2;
            }
            xy1[4] = a1 * i4xy;
            xy2[4] = a2 * i4xy;
            if (o >= z0) {
                const float fzz1 = invdzz * central2(cq1[4], cq1[3], cq1[5], cq1[2], cq1[6], cq1[1], cq1[7], cq1[0],
                                                     cq1[8]);
                const float fzz2 = invdzz * central2(cq2[4], cq2[3], cq2[5], cq2[2], cq2[6], cq2[1], cq2[7], cq2[0],
                                                     cq2[8]);
                const float vx = pw[4], vy = pw[5], vz = pw[6];
                const float rx2 = vx * vx, ry2 = vy * vy, rz2 = vz * vz;
                const float rxy = 2.0f * vx * vy, ryz = 2.0f * vy * vz, rzx = 2.0f * vz * vx;
                const float f1zx = qx1[0] * i4zx, f2zx = qx2[0] * i4zx, f1yz = qy1[0] * i4yz, f2yz = qy2[0] * i4yz;
                const float V1 = rx2 * xx1[0] + ry2 * yy1[0] + rz2 * fzz1 + rxy * xy1[0] + rzx * f1zx + ryz * f1yz;
                const float H1 = xx1[0] + yy1[0] + fzz1 - V1;
                const float V2 = rx2 * xx2[0] + ry2 * yy2[0] + rz2 * fzz2 + rxy * xy2[0] + rzx * f2zx + ryz * f2yz;
                const float H2 = xx2[0] + yy2[0] + fzz2 - V2;
                const float mm1 = pw[0], mm2 = pw[1], mm3 = pw[2], mm4 = pw[3];
                const float r3 = mm1 * H1 + mm4 * V1 + mm2 * V2;
                const float r4 = mm2 * H1 + mm3 * V2 + mm4 * H2;
                if (MODE == 2) {
                    un = r3;
                    wn = r4;
                } else if (MODE == 1) {
                    const float a = fminf(axy, abz[o].a);
                    const float bb = fminf(bxy, abz[o].b);
                    un = a * pw[7] + bb * r3 * dt;
                    wn = a * pw[8] + bb * r4 * dt;
                } else {
                    un = pw[7] + r3 * dt;
                    wn = pw[8] + r4 * dt;
                }
            }
        }
        load_pw(o + 1);
        if (active && o >= z0) {
            if (MODE == 2) {
                const long oi = ((long)(o - izbeg) * oldy + (iy - iybeg)) * oldx + (ix - ixbeg);
                u[oi] = un;
                w[oi] = wn;
            } else {
                const long i = (long)o * sz + col;
                u[i] = un;
                w[i] = wn;
            }
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            cq1[k] = cq1[k + 1];
            cq2[k] = cq2[k + 1];
            qx1[k] = qx1[k + 1];
            qx2[k] = qx2[k + 1];
            qy1[k] = qy1[k + 1];
            qy2[k] = qy2[k + 1];
        }
        qx1[8] = qx2[8] = qy1[8] = qy2[8] = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            xx1[k] = xx1[k + 1];
            xx2[k] = xx2[k + 1];
            yy1[k] = yy1[k + 1];
            yy2[k] = yy2[k + 1];
            xy1[k] = xy1[k + 1];
            xy2[k] = xy2[k + 1];
        }
        __syncthreads();
    }
}

// The transposed operator with the plane pipeline: the 9 input fields of the (H x W) tile of plane p + 2 are loaded
// as float4 quads (one per thread) during plane p and turned into T_x, T_y, M_xy, M_zx, M_yz (and T_z of every column)
// in the other LDS buffer during plane p + 1; u, w of the next output plane are loaded one plane ahead; one barrier
// per plane. Needs quads_fit (the launcher falls back).
template <int BY, bool SPONGE>
__global__ __launch_bounds__(BX* BY) void k_t1_adj_q(
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ f1, const float* __restrict__ f2,
    const float* __restrict__ m1, const float* __restrict__ m2, const float* __restrict__ m3,
    const float* __restrict__ m4, const float* __restrict__ m5, const float* __restrict__ m6,
    const float* __restrict__ m7, const ab2* __restrict__ abx, const ab2* __restrict__ aby,
    const ab2* __restrict__ abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx,
    long ldy, float dt, float invdxx, float invdyy, float invdzz, float i4xy, float i4yz, float i4zx, int bz)
{
    constexpr int W = BX + 2 * R, H = BY + 2 * R, QPR = W / 4, NQ = QPR * H;
    static_assert(NQ <= BX * BY, "one tile quad per thread");
    // the two fields of each temporary interleaved: one 8-byte LDS read per stencil tap gives both (2.75 -> 2.73 ms
    // at 512^3, bit-identical)
    __shared__ __align__(16) float2 ltx[2][BY][W];   // T_x (x halo)
    __shared__ __align__(16) float2 lty[2][H][BX];   // T_y (y halo)
    __shared__ __align__(16) float2 lmxy[2][H][W];   // M_xy
    __shared__ __align__(16) float2 lmzx[2][BY][W];  // M_zx (x halo)
    __shared__ __align__(16) float2 lmyz[2][H][BX];  // M_yz (y halo)
    __shared__ __align__(16) float2 ltz[2][BY][BX];  // T_z (centre)
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
    const int p0 = z0 - R, pend = z1 + R;

    const bool qhas = tid < NQ;
    const int qr = tid / QPR, ql = 4 * (tid - qr * QPR);
    const int qkeep = qhas && qr < ylim ? min(max(xlim - ql, 0), 4) : 0;
    const long qoff = (long)(y0 - R + qr) * ldx + (x0 - R + ql);
    const bool qrow = qr >= R && qr < R + BY, qcol = ql >= R && ql < R + BX;
    float4 v1, v2, vm1, vm2, vm3, vm4, vm5, vm6, vm7;
    auto fetch = [&](int p) {
        if (qkeep > 0) {
            const long g = (long)p * sz + qoff;
            v1 = *reinterpret_cast<const float4*>(f1 + g);
            v2 = *reinterpret_cast<const float4*>(f2 + g);
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
        const float* M1 = &vm1.x;
        const float* M2 = &vm2.x;
        const float* M3 = &vm3.x;
        const float* M4 = &vm4.x;
        const float* RX = &vm5.x;
        const float* RY = &vm6.x;
        const float* RZ = &vm7.x;
        float vx1[4], vx2[4], vy1[4], vy2[4], vxy1[4], vxy2[4], vzx1[4], vzx2[4], vyz1[4], vyz2[4], vz1[4], vz2[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            vx1[e] = vx2[e] = vy1[e] = vy2[e] = vxy1[e] = vxy2[e] = 0.0f;
            vzx1[e] = vzx2[e] = vyz1[e] = vyz2[e] = vz1[e] = vz2[e] = 0.0f;
            if (e < qkeep) {
                const float a = A1[e], c = A2[e], mm1 = M1[e], mm2 = M2[e], mm3 = M3[e], mm4 = M4[e];
                const float lap1 = mm1 * a + mm2 * c, hz1 = (-mm1 + mm4) * a - mm2 * c;
                const float lap2 = mm4 * c, hz2 = mm2 * a + (-mm4 + mm3) * c;
                const float rx = RX[e], ry = RY[e], rz = RZ[e];
                vx1[e] = lap1 + rx * rx * hz1;
                vx2[e] = lap2 + rx * rx * hz2;
                vy1[e] = lap1 + ry * ry * hz1;
                vy2[e] = lap2 + ry * ry * hz2;
                vxy1[e] = 2.0f * rx * ry * hz1;
                vxy2[e] = 2.0f * rx * ry * hz2;
                vzx1[e] = 2.0f * rz * rx * hz1;
                vzx2[e] = 2.0f * rz * rx * hz2;
                vyz1[e] = 2.0f * ry * rz * hz1;
                vyz2[e] = 2.0f * ry * rz * hz2;
                vz1[e] = mm1 * a + mm2 * c + rz * rz * ((-mm1 + mm4) * a - mm2 * c);
                vz2[e] = mm4 * c + rz * rz * (mm2 * a + (-mm4 + mm3) * c);
            }
        }
        auto st2 = [](float2* dst, const float* a, const float* b) { // 4 points of two fields, interleaved
            *reinterpret_cast<float4*>(dst) = make_float4(a[0], b[0], a[1], b[1]);
            *reinterpret_cast<float4*>(dst + 2) = make_float4(a[2], b[2], a[3], b[3]);
        };
        st2(&lmxy[bf][qr][ql], vxy1, vxy2);
        if (qrow) {
            st2(&ltx[bf][qr - R][ql], vx1, vx2);
            st2(&lmzx[bf][qr - R][ql], vzx1, vzx2);
        }
        if (qcol) {
            st2(&lty[bf][qr][ql - R], vy1, vy2);
            st2(&lmyz[bf][qr][ql - R], vyz1, vyz2);
        }
        if (qrow && qcol)
            st2(&ltz[bf][qr - R][ql - R], vz1, vz2);
    };
    float axy = 1.0f, bxy = 1.0f;
    if (SPONGE && active) {
        axy = fminf(abx[ix].a, aby[iy].a);
        bxy = fminf(abx[ix].b, aby[iy].b);
    }
    float pu = 0.0f, pwv = 0.0f; // u, w of the output plane being finalised
    auto load_pw = [&](int o) {
        if (active && o >= z0 && o < z1) {
            pu = u[(long)o * sz + col];
            pwv = w[(long)o * sz + col];
        }
    };

    float cq1[9], cq2[9];                                  // T_z(p - 8 + k)
    float cx1[5], cx2[5], cy1[5], cy2[5], xy1[5], xy2[5]; // central2_x(T_x), central2_y(T_y), X_xy(M_xy) at p-4+k
    float q1[9], q2[9];                                    // X_zx + X_yz partial sums
#pragma unroll
    for (int k = 0; k < 9; ++k)
        cq1[k] = cq2[k] = q1[k] = q2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        cx1[k] = cx2[k] = cy1[k] = cy2[k] = xy1[k] = xy2[k] = 0.0f;

    load_pw(p0 - R + 1);
    fetch(p0);
    commit(0);
    if (p0 + 1 < pend)
        fetch(p0 + 1);
    __syncthreads();
    for (int p = p0; p < pend; ++p) {
        const int o = p - R, bf = (p - p0) & 1;
        if (p + 1 < pend)
            commit(bf ^ 1);
        if (p + 2 < pend)
            fetch(p + 2);
        float un = 0.0f, wn = 0.0f;
        if (active) {
            const float2(*TX)[W] = ltx[bf];
            const float2(*TY)[BX] = lty[bf];
            const float2(*MXY)[W] = lmxy[bf];
            const float2(*MZX)[W] = lmzx[bf];
            const float2(*MYZ)[BX] = lmyz[bf];
            const float2 cz = ltz[bf][ty][tx];
            cq1[8] = cz.x;
            cq2[8] = cz.y;
            const int cx = tx + R, cy = ty + R;
            float2 vx[9], vy[9]; // taps at offsets 0, -1, +1, ..., -4, +4 (the central2 argument order)
#pragma unroll
            for (int n = 0; n <= R; ++n) {
                vx[n == 0 ? 0 : 2 * n - 1] = TX[ty][cx - n];
                vy[n == 0 ? 0 : 2 * n - 1] = TY[cy - n][tx];
                if (n > 0) {
                    vx[2 * n] = TX[ty][cx + n];
                    vy[2 * n] = TY[cy + n][tx];
                }
            }
            cx1[4] = invdxx * central2(vx[0].x, vx[1].x, vx[2].x, vx[3].x, vx[4].x, vx[5].x, vx[6].x, vx[7].x, vx[8].x);
            cx2[4] = invdxx * central2(vx[0].y, vx[1].y, vx[2].y, vx[3].y, vx[4].y, vx[5].y, vx[6].y, vx[7].y, vx[8].y);
            cy1[4] = invdyy * central2(vy[0].x, vy[1].x, vy[2].x, vy[3].x, vy[4].x, vy[5].x, vy[6].x, vy[7].x, vy[8].x);
            cy2[4] = invdyy * central2(vy[0].y, vy[1].y, vy[2].y, vy[3].y, vy[4].y, vy[5].y, vy[6].y, vy[7].y, vy[8].y);
            float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float e = ecen(n);
                const float2 pp = MXY[cy + n][cx + n], pm = MXY[cy + n][cx - n];
                const float2 mm = MXY[cy - n][cx - n], mp = MXY[cy - n][cx + n];
                a1 += e * (pp.x - pm.x + mm.x - mp.x);
                a2 += e * (pp.y - pm.y + mm.y - mp.y);
                const float2 zxp = MZX[ty][cx + n], zxm = MZX[ty][cx - n];
                const float2 yzp = MYZ[cy + n][tx], yzm = MYZ[cy - n][tx];
                const float d1 = i4zx * (zxp.x - zxm.x) + i4yz * (yzp.x - yzm.x);
                const float d2 = i4zx * (zxp.y - zxm.y) + i4yz * (yzp.y - yzm.y);
                q1[4 - n] += e * d1;
                q1[4 + n] -= e * d1;
                q2[4 - n] += e * d2;
                q2[4 + n] -= e * d2;
            }
            xy1[4] = a1 * i4xy;
            xy2[4] = a2 * i4xy;
            if (o >= z0) {
                const float cz1 = invdzz * central2(cq1[4], cq1[3], cq1[5], cq1[2], cq1[6], cq1[1], cq1[7], cq1[0],
                                                    cq1[8]);
                const float cz2 = invdzz * central2(cq2[4], cq2[3], cq2[5], cq2[2], cq2[6], cq2[1], cq2[7], cq2[0],
                                                    cq2[8]);
                const float r3 = cx1[0] + cy1[0] + cz1 + xy1[0] + q1[0];
                const float r4 = cx2[0] + cy2[0] + cz2 + xy2[0] + q2[0];
                if (SPONGE) {
                    const float a = fminf(axy, abz[o].a);
                    const float bb = fminf(bxy, abz[o].b);
                    un = a * pu + bb * r3 * dt;
                    wn = a * pwv + bb * r4 * dt;
                } else {
                    un = pu + r3 * dt;
                    wn = pwv + r4 * dt;
                }
            }
        }
        load_pw(o + 1);
        if (active && o >= z0) {
            const long i = (long)o * sz + col;
            u[i] = un;
            w[i] = wn;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            cq1[k] = cq1[k + 1];
            cq2[k] = cq2[k + 1];
            q1[k] = q1[k + 1];
            q2[k] = q2[k + 1];
        }
        q1[8] = q2[8] = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cx1[k] = cx1[k + 1];
            cx2[k] = cx2[k + 1];
            cy1[k] = cy1[k + 1];
            cy2[k] = cy2[k + 1];
            xy1[k] = xy1[k + 1];
            xy2[k] = xy2[k + 1];
        }
        __syncthreads();
    }
}

} // namespace tti

using namespace tti;

extern "C" {

int
fwd_main_loop_2_1_inner_gpu_launch(float* field3, float* field4, const float* field1, const float* field2,
                                   const float* model1, const float* model2, const float* model3, const float* model4,
                                   const float* model5, const float* model6, const float* model7, const void* ABx,
                                   const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend,
                                   int izbeg, int izend, int ldimx_in, int ldimy_in, int ldimz_in, int ldimx_out,
                                   int ldimy_out, int ldimz_out, float dt, double dx, double dy, double dz, int order,
                                   void* stream, int calc_snap, int use_sponge, int simple)
{
    (void)ldimz_in, (void)ldimz_out, (void)simple;
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!field3 || !field4 || !field1 || !field2 || !model1 || !model2 || !model3 || !model4 || !model5 || !model6 ||
        !model7)
        return TTI_EINVAL;
    if (calc_snap && (ldimx_out < ixend - ixbeg + 1 || ldimy_out < iyend - iybeg + 1))
        return TTI_EINVAL;
    // the simple reference kernel always applies the sponge form (with the coefficient arrays it is given)
    const bool sponge = !calc_snap && (use_sponge || simple);
    if (sponge && (!ABx || !ABy || !ABz))
        return TTI_EINVAL;
    if (order == 16)
        return order_t1_fwd(order, field3, field4, field1, field2, model1, model2, model3, model4, model5, model6,
                            model7, ABx, ABy, ABz, ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx_in, ldimy_in,
                            ldimx_out, ldimy_out, dt, dx, dy, dz, calc_snap ? 2 : sponge ? 1 : 0, stream);
    if (!ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx_in, ldimy_in))
        return TTI_EINVAL;
    const float invdxx = (float)(1.0 / (dx * dx)), invdyy = (float)(1.0 / (dy * dy)), invdzz = (float)(1.0 / (dz * dz));
    const float i4xy = (float)(1.0 / (4 * dx * dy)), i4yz = (float)(1.0 / (4 * dy * dz)), i4zx = (float)(1.0 / (4 * dx * dz));
    const int bz = stencil_bz(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    const dim3 grid = stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz);
    hipStream_t s = (hipStream_t)stream;
    if (quads_fit(ixbeg, ldimx_in, { field1, field2, model1, model2, model3, model4, model5, model6, model7 })) {
        auto kern = calc_snap ? k_t1_fwd_q<BYS, 2> : sponge ? k_t1_fwd_q<BYS, 1> : k_t1_fwd_q<BYS, 0>;
        kern<<<grid, dim3(BX, BYS), 0, s>>>(field3, field4, field1, field2, model1, model2, model3, model4, model5,
                                            model6, model7, sponge ? (const ab2*)ABx : nullptr,
                                            sponge ? (const ab2*)ABy : nullptr, sponge ? (const ab2*)ABz : nullptr,
                                            ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx_in, (long)ldimy_in,
                                            calc_snap ? (long)ldimx_out : 0L, calc_snap ? (long)ldimy_out : 0L, dt,
                                            invdxx, invdyy, invdzz, i4xy, i4yz, i4zx, bz);
    } else if (calc_snap)
        k_t1_fwd<BYS, 2><<<grid, dim3(BX, BYS), 0, s>>>(field3, field4, field1, field2, model1, model2, model3, model4,
                                                      model5, model6, model7, nullptr, nullptr, nullptr, ixbeg, ixend,
                                                      iybeg, iyend, izbeg, izend, (long)ldimx_in, (long)ldimy_in,
                                                      (long)ldimx_out, (long)ldimy_out, dt, invdxx, invdyy, invdzz,
                                                      i4xy, i4yz, i4zx, bz);
    else if (sponge)
        k_t1_fwd<BYS, 1><<<grid, dim3(BX, BYS), 0, s>>>(field3, field4, field1, field2, model1, model2, model3, model4,
                                                      model5, model6, model7, (const ab2*)ABx, (const ab2*)ABy,
                                                      (const ab2*)ABz, ixbeg, ixend, iybeg, iyend, izbeg, izend,
                                                      (long)ldimx_in, (long)ldimy_in, 0L, 0L, dt, invdxx, invdyy,
                                                      invdzz, i4xy, i4yz, i4zx, bz);
    else
        k_t1_fwd<BYS, 0><<<grid, dim3(BX, BYS), 0, s>>>(field3, field4, field1, field2, model1, model2, model3, model4,
                                                      model5, model6, model7, nullptr, nullptr, nullptr, ixbeg, ixend,
                                                      iybeg, iyend, izbeg, izend, (long)ldimx_in, (long)ldimy_in, 0L,
                                                      0L, dt, invdxx, invdyy, invdzz, i4xy, i4yz, i4zx, bz);
    return launch_status();
}

int
adj_main_loop_2_inner_gpu_launch(float* field3, float* field4, const float* field1, const float* field2,
                                 const float* model1, const float* model2, const float* model3, const float* model4,
                                 const float* model5, const float* model6, const float* model7, const void* ABx,
                                 const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend,
                                 int izbeg, int izend, int ldimx, int ldimy, int ldimz, float dt, double dx, double dy,
                                 double dz, int order, void* stream, int use_sponge, int simple)
{
    (void)ldimz;
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!field3 || !field4 || !field1 || !field2 || !model1 || !model2 || !model3 || !model4 || !model5 || !model6 ||
        !model7)
        return TTI_EINVAL;
    const bool sponge = use_sponge || simple;
    if (sponge && (!ABx || !ABy || !ABz))
        return TTI_EINVAL;
    if (order == 16)
        return order_t1_adj(order, field3, field4, field1, field2, model1, model2, model3, model4, model5, model6,
                            model7, ABx, ABy, ABz, ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, dt, dx, dy,
                            dz, sponge ? 1 : 0, stream);
    if (!ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy))
        return TTI_EINVAL;
    const float invdxx = (float)(1.0 / (dx * dx)), invdyy = (float)(1.0 / (dy * dy)), invdzz = (float)(1.0 / (dz * dz));
    const float i4xy = (float)(1.0 / (4 * dx * dy)), i4yz = (float)(1.0 / (4 * dy * dz)), i4zx = (float)(1.0 / (4 * dx * dz));
    const int bz = stencil_bz(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    const dim3 grid = stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz);
    hipStream_t s = (hipStream_t)stream;
    if (quads_fit(ixbeg, ldimx, { field1, field2, model1, model2, model3, model4, model5, model6, model7 })) {
        auto kern = sponge ? k_t1_adj_q<BYS, true> : k_t1_adj_q<BYS, false>;
        kern<<<grid, dim3(BX, BYS), 0, s>>>(field3, field4, field1, field2, model1, model2, model3, model4, model5,
                                            model6, model7, sponge ? (const ab2*)ABx : nullptr,
                                            sponge ? (const ab2*)ABy : nullptr, sponge ? (const ab2*)ABz : nullptr,
                                            ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy, dt,
                                            invdxx, invdyy, invdzz, i4xy, i4yz, i4zx, bz);
    } else if (sponge)
        k_t1_adj<BYS, true><<<grid, dim3(BX, BYS), 0, s>>>(field3, field4, field1, field2, model1, model2, model3,
                                                         model4, model5, model6, model7, (const ab2*)ABx,
                                                         (const ab2*)ABy, (const ab2*)ABz, ixbeg, ixend, iybeg, iyend,
                                                         izbeg, izend, (long)ldimx, (long)ldimy, dt, invdxx, invdyy,
                                                         invdzz, i4xy, i4yz, i4zx, bz);
    else
        k_t1_adj<BYS, false><<<grid, dim3(BX, BYS), 0, s>>>(field3, field4, field1, field2, model1, model2, model3,
                                                          model4, model5, model6, model7, nullptr, nullptr, nullptr,
                                                          ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx,
                                                          (long)ldimy, dt, invdxx, invdyy, invdzz, i4xy, i4yz, i4zx,
                                                          bz);
    return launch_status();
}

int update_Vp_gradient(const float*, const float*, float*, const float*, const float*, int, int, int, int, int, int,
                       int, int, int, int, void*);

int
launch_update_grad_loop_kernel(const float* snap_x, const float* snap_z, float* grad_Vp, const float* adj_x,
                               const float* adj_z, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                               int snap_ldimx, int snap_ldimy, int ldimx, int ldimy, void* stream)
{
    // same imaging condition as the type2 family (kernels/kernel_type1_gpu_grad.cu:5-22)
    return update_Vp_gradient(snap_x, snap_z, grad_Vp, adj_x, adj_z, ixbeg, ixend, iybeg, iyend, izbeg, izend,
                              snap_ldimx, snap_ldimy, ldimx, ldimy, stream);
}

} // extern "C"
