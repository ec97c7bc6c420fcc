// Constant-density TTI operator (the "type1" family) for gfx950: forward, adjoint and imaging condition.
//
// Central second derivatives F_aa = (1/da^2)(e0 f + sum_n e_n (f(-n) + f(+n))) replace the staggered D+ b D- pair of
// the variable-density operator; the rotated mixed derivatives and the V/H model combination are the same
// (kernels/kernel_type1_gpu_fwd.cu:9-71, :225-437; kernels/kernel_type1_gpu_adj.cu:9-51, :235-378;
// kernels/helper_kernels_gpu.h:120-418). Same 2.5-D structure as the type2 kernels: 64 x 8 tiles in LDS, a z-march
// with the z stencil in a 9-deep register queue, and plane p = o + 4 absorbed into the x-z / y-z partial sums.
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
    const int x0 = ixbeg + blockIdx.x * BX, y0 = iybeg + blockIdx.y * BY, z0 = izbeg + blockIdx.z * bz;
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
    const int x0 = ixbeg + blockIdx.x * BX, y0 = iybeg + blockIdx.y * BY, z0 = izbeg + blockIdx.z * bz;
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
    if (order != 8)
        return TTI_EORDER;
    if (!field3 || !field4 || !field1 || !field2 || !model1 || !model2 || !model3 || !model4 || !model5 || !model6 ||
        !model7 || !ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx_in, ldimy_in))
        return TTI_EINVAL;
    if (calc_snap && (ldimx_out < ixend - ixbeg + 1 || ldimy_out < iyend - iybeg + 1))
        return TTI_EINVAL;
    // the simple reference kernel always applies the sponge form (with the coefficient arrays it is given)
    const bool sponge = !calc_snap && (use_sponge || simple);
    if (sponge && (!ABx || !ABy || !ABz))
        return TTI_EINVAL;
    const float invdxx = (float)(1.0 / (dx * dx)), invdyy = (float)(1.0 / (dy * dy)), invdzz = (float)(1.0 / (dz * dz));
    const float i4xy = (float)(1.0 / (4 * dx * dy)), i4yz = (float)(1.0 / (4 * dy * dz)), i4zx = (float)(1.0 / (4 * dx * dz));
    const int bz = BZ_DEFAULT;
    const dim3 grid = stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz);
    hipStream_t s = (hipStream_t)stream;
    if (calc_snap)
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
    if (order != 8)
        return TTI_EORDER;
    if (!field3 || !field4 || !field1 || !field2 || !model1 || !model2 || !model3 || !model4 || !model5 || !model6 ||
        !model7 || !ranges_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy))
        return TTI_EINVAL;
    const bool sponge = use_sponge || simple;
    if (sponge && (!ABx || !ABy || !ABz))
        return TTI_EINVAL;
    const float invdxx = (float)(1.0 / (dx * dx)), invdyy = (float)(1.0 / (dy * dy)), invdzz = (float)(1.0 / (dz * dz));
    const float i4xy = (float)(1.0 / (4 * dx * dy)), i4yz = (float)(1.0 / (4 * dy * dz)), i4zx = (float)(1.0 / (4 * dx * dz));
    const int bz = BZ_DEFAULT;
    const dim3 grid = stencil_grid(ixbeg, ixend, iybeg, iyend, izbeg, izend, bz);
    hipStream_t s = (hipStream_t)stream;
    if (sponge)
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
