// Fused single-pass TTI time steps for gfx950: one kernel per forward step and one per adjoint step.
//
// The reference runs a forward step as loop2 -> derivatives -> update (three passes, each a full HBM sweep,
// simulator2/simulator2SetUp.cu:530-599). Here one kernel reads sigma_{n+1} (2 fields), u_n (2), the 8 model words,
// and writes u_{n+1} and sigma_{n+2}: 16 words = 64 B per grid-point update. The time update of sigma is rotated
// into the same kernel (sigma_{n+2} = A sigma_{n+1} + B dt u_{n+1} is pointwise), so sigma ping-pongs between two
// buffers and u is updated in place.
//
// The adjoint step (simulator2/simulator2SetUp.cu:662-737: adjoint derivatives -> update -> loop2) is naturally a
// single pass: u' = u + dt L^T(sigma) (or the sponge form), sigma' = A sigma + B dt u'.  The zero-lag imaging
// condition g += u' * u_snap + w' * w_snap (kernels/kernel_type2_small.cu:42-59) is fused into it, accumulating in
// FP64, with the snapshot read in FP32, FP16 (per-64-point power-of-two scale) or BF16.
//
// Ghost semantics are those of the reference single-volume case: derivative fields are zero outside the interior
// (the reference writes them on the interior only and never fills their outer ghosts) and sigma ghosts hold what
// the caller stored there.
//
// Block = 64 x BY threads, one output column per thread, marching over a chunk of z planes. Per plane p the block
// stages sigma (halo 8) and b (halo 4) tiles in LDS, builds the staggered derivatives d = b D- sigma of plane p on
// the tile plus a 4-point rim, and absorbs plane p into register queues; output plane o = p - 4 is finalised with
// the z part coming from a per-thread column stream 7 planes ahead.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "common.h"
#include "tti_fused.h"

#include <initializer_list>

namespace tti {

namespace fused {

constexpr int BX = 64;
constexpr int BY = 8;
constexpr int TW = BX + 16, TH = BY + 16; // sigma tile with halo 8
constexpr int BW = BX + 8, BH = BY + 8;   // halo-4 tile

struct Snap
{
    int mode; // TTI_SNAP_NONE / FP32 / FP16 / BF16
    void* a;
    void* b;
    float* sa;
    float* sb;
    long ldx; // row length of the snapshot arrays (multiple of 64)
    long ldy;
};

struct Box
{
    int ixbeg, ixend, iybeg, iyend, izbeg, izend;
    long ldx, ldy;
};

__device__ __forceinline__ bool
interior(const Box& g, int x, int y, int z)
{
    return x >= g.ixbeg && x <= g.ixend && y >= g.iybeg && y <= g.iyend && z >= g.izbeg && z <= g.izend;
}

// valid (allocated) region: interior plus a 4-point ghost rim
__device__ __forceinline__ bool
allocated(const Box& g, int x, int y, int z)
{
    return x >= g.ixbeg - R && x <= g.ixend + R && y >= g.iybeg - R && y <= g.iyend + R && z >= g.izbeg - R &&
           z <= g.izend + R;
}

// XCD-aware tile decode: consecutive hardware blocks go round-robin over the 8 XCDs; each XCD gets a contiguous range
// of tiles (x fastest, then y, then z chunk) so neighbouring tiles share halo lines in the same L2.
__device__ __forceinline__ bool
decode_tile(int ntx, int nty, int ntz, int& bxi, int& byi, int& bzi)
{
    const int nt = ntx * nty * ntz;
    const int bid = blockIdx.x;
    const int per = (nt + 7) / 8;
    const int t = (bid % 8) * per + bid / 8;
    if (t >= nt)
        return false;
    bxi = t % ntx;
    byi = (t / ntx) % nty;
    bzi = t / (ntx * nty);
    return true;
}

__device__ __forceinline__ float
wave_max(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
        v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}

// Stores one 64-point row segment of a snapshot field (all 64 lanes of the wave call this).
__device__ __forceinline__ void
snap_store(int mode, void* arr, float* scales, long idx, long sidx, bool act, float v)
{
    if (mode == TTI_SNAP_FP32) {
        if (act)
            reinterpret_cast<float*>(arr)[idx] = v;
    } else if (mode == TTI_SNAP_FP16) {
        const float m = wave_max(act ? fabsf(v) : 0.0f);
        int e = 0;
        frexpf(m, &e); // m < 2^e
        if (m == 0.0f)
            e = 0;
        if (act)
            reinterpret_cast<__half*>(arr)[idx] = __float2half(ldexpf(v, 14 - e));
        if (threadIdx.x == 0 && sidx >= 0)
            scales[sidx] = ldexpf(1.0f, e - 14);
    } else if (mode == TTI_SNAP_BF16) {
        if (act)
            reinterpret_cast<__hip_bfloat16*>(arr)[idx] = __float2bfloat16(v);
    }
}

__device__ __forceinline__ float
snap_load(int mode, const void* arr, const float* scales, long idx, long sidx)
{
    if (mode == TTI_SNAP_FP32)
        return reinterpret_cast<const float*>(arr)[idx];
    if (mode == TTI_SNAP_FP16)
        return __half2float(reinterpret_cast<const __half*>(arr)[idx]) * scales[sidx];
    if (mode == TTI_SNAP_BF16)
        return __bfloat162float(reinterpret_cast<const __hip_bfloat16*>(arr)[idx]);
    return 0.0f;
}

// Masked float4 load of 4 consecutive x points starting at gx (gx = 0 mod 4). Points outside the allocated
// x range [ixbeg-4, ixend+4] read as zero; quads wholly outside are not loaded. With ixbeg = 0 mod 4 and
// ldx = 0 mod 4 (checked by the launchers) a quad that overlaps the allocated range lies inside [0, ldx).
__device__ __forceinline__ float4
load_quad(const float* __restrict__ base, long idx, int gx, const Box& g, bool rowok)
{
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (rowok && gx + 3 >= g.ixbeg - R && gx <= g.ixend + R) {
        v = *reinterpret_cast<const float4*>(base + idx);
        const int lim = g.ixend + R - gx; // last valid element offset within the quad
        if (lim < 3) {
            if (lim < 1)
                v.y = 0.0f;
            if (lim < 2)
                v.z = 0.0f;
            v.w = 0.0f;
        }
    }
    return v;
}

// Quad mappings of the two tile shapes onto the 512 threads of a block.
constexpr int TQ_PER_ROW = TW / 4, TQ = TQ_PER_ROW * TH; // 20 x 24 = 480 quads
constexpr int BQ_PER_ROW = BW / 4, BQ = BQ_PER_ROW * BH; // 18 x 16 = 288 quads
static_assert(TQ <= BX * BY && BQ <= BX * BY, "one quad per thread");

// ------------------------------------------------------------------------------------------------------------------
// Forward step: u += dt L(sigma) (+ source), sigma_out = A sigma + B dt u_new; optional snapshot of u_new.
// Same arithmetic as kernels/kernel_type2_fwd_drv.cu:148-362 + kernels/kernel_type2_fwd.cu:264-540 +
// kernels/kernel_base.cu:82-104, evaluated in one sweep.
//
// Pipeline per output plane o (tile plane p = o + 4): the tile of plane p is already in LDS buffer `cur`;
//   A  issue the column loads (sigma(o+7), b(o+4)) and the pointwise loads of plane o (models, u, w)
//   B  staggered derivatives d_x, d_y of plane p on the tile rim -> LDS, barrier
//   C  D+ of d_x, d_y, the x-y mixed term and the z-mixed partial sums of plane p (registers)
//   D  d_z(o+4) from the column queue, finalise plane o
//   E  store the prefetched tile of plane p+1 into buffer 1-cur, prefetch plane p+2 into registers, barrier
// ------------------------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BX* BY) void k_fused_fwd(
    const float* __restrict__ s1, const float* __restrict__ s2, float* __restrict__ s1o, float* __restrict__ s2o,
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ m1, const float* __restrict__ m2,
    const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ vrx,
    const float* __restrict__ vry, const float* __restrict__ vrz, const float* __restrict__ b,
    const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, Box g, float dt,
    float invdx, float invdy, float invdz, float i4xy, float i4yz, float i4zx, long src_idx, float src_val, Snap snap,
    int bz, int ntx, int nty, int ntz)
{
    __shared__ __align__(16) float T1[2][TH][TW];
    __shared__ __align__(16) float T2[2][TH][TW]; // sigma, halo 8: (x0-8+lx, y0-8+ly)
    __shared__ __align__(16) float Bt[2][BH][BW]; // b, halo 4: (x0-4+lx, y0-4+ly)
    __shared__ float X1[BY][BW], X2[BY][BW];      // d_x at (x0-4+j, y0+ty)
    __shared__ float Y1[BH][BX], Y2[BH][BX];      // d_y at (x0+tx, y0-4+j)

    int bxi, byi, bzi;
    if (!decode_tile(ntx, nty, ntz, bxi, byi, bzi))
        return;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    const int x0 = g.ixbeg + bxi * BX, y0 = g.iybeg + byi * BY, z0 = g.izbeg + bzi * bz;
    const int z1 = min(z0 + bz, g.izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= g.ixend && iy <= g.iyend;
    const long sz = g.ldx * g.ldy;
    const long col = (long)iy * g.ldx + ix;
    const bool sponge = abx != nullptr;
    float axy = 1.0f, bxy = 1.0f;
    if (sponge && active) {
        axy = fminf(abx[ix].a, aby[iy].a);
        bxy = fminf(abx[ix].b, aby[iy].b);
    }
    // quad maps
    const bool tq_on = tid < TQ, bq_on = tid < BQ;
    const int tq_row = tid / TQ_PER_ROW, tq_lx = 4 * (tid - tq_row * TQ_PER_ROW);
    const int bq_row = tid / BQ_PER_ROW, bq_lx = 4 * (tid - bq_row * BQ_PER_ROW);
    const int tq_gx = x0 - 8 + tq_lx, tq_gy = y0 - 8 + tq_row;
    const int bq_gx = x0 - 4 + bq_lx, bq_gy = y0 - 4 + bq_row;
    const bool tq_ok = tq_on && tq_gy >= g.iybeg - R && tq_gy <= g.iyend + R;
    const bool bq_ok = bq_on && bq_gy >= g.iybeg - R && bq_gy <= g.iyend + R;
    const long tq_off = (long)tq_gy * g.ldx + tq_gx, bq_off = (long)bq_gy * g.ldx + bq_gx;
    float4 pt1, pt2, pb;
    auto fetch = [&](int pp) {
        const long org = (long)pp * sz;
        pt1 = load_quad(s1, org + tq_off, tq_gx, g, tq_ok);
        pt2 = load_quad(s2, org + tq_off, tq_gx, g, tq_ok);
        pb = load_quad(b, org + bq_off, bq_gx, g, bq_ok);
    };
    auto commit = [&](int buf) {
        if (tq_on) {
            *reinterpret_cast<float4*>(&T1[buf][tq_row][tq_lx]) = pt1;
            *reinterpret_cast<float4*>(&T2[buf][tq_row][tq_lx]) = pt2;
        }
        if (bq_on)
            *reinterpret_cast<float4*>(&Bt[buf][bq_row][bq_lx]) = pb;
    };

    float sq1[8], sq2[8]; // sigma(o + k)
    float dz1[8], dz2[8]; // d_z(o - 3 + k)
    float xx1[5], xx2[5]; // D+_x d_x (o + k), unscaled by rho
    float yy1[5], yy2[5]; // D+_y d_y
    float xy1[5], xy2[5]; // X_xy sigma
    float qx1[9], qx2[9]; // z-x mixed partial sums, q[j] <-> output p - 4 + j
    float qy1[9], qy2[9]; // y-z mixed partial sums
#pragma unroll
    for (int k = 0; k < 8; ++k)
        dz1[k] = dz2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        xx1[k] = xx2[k] = yy1[k] = yy2[k] = xy1[k] = xy2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k)
        qx1[k] = qx2[k] = qy1[k] = qy2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const int q = z0 - 8 + k;
        const bool ok = active && q >= g.izbeg - R && q <= g.izend + R;
        sq1[k] = ok ? s1[(long)q * sz + col] : 0.0f;
        sq2[k] = ok ? s2[(long)q * sz + col] : 0.0f;
    }
    int cur = 0;
    fetch(z0 - R);
    commit(0);
    fetch(z0 - R + 1);
    __syncthreads();

    for (int o = z0 - 8; o < z1; ++o) {
        const int p = o + R;
        const bool emit = o >= z0;
        // ---- A: column and pointwise loads
        {
            const int q = o + 7;
            const bool ok = active && q >= g.izbeg - R && q <= g.izend + R;
            sq1[7] = ok ? s1[(long)q * sz + col] : 0.0f;
            sq2[7] = ok ? s2[(long)q * sz + col] : 0.0f;
        }
        const int qd = o + 4;
        const bool dzok = active && qd >= g.izbeg && qd <= g.izend;
        const float bd = dzok ? b[(long)qd * sz + col] : 0.0f;
        const bool fin = emit && active;
        const long io = (long)o * sz + col;
        float mm1 = 0, mm2 = 0, mm3 = 0, mm4 = 0, vx = 0, vy = 0, vz = 0, bo = 1, uo = 0, wo = 0;
        if (fin) {
            mm1 = m1[io];
            mm2 = m2[io];
            mm3 = m3[io];
            mm4 = m4[io];
            vx = vrx[io];
            vy = vry[io];
            vz = vrz[io];
            bo = b[io];
            uo = u[io];
            wo = w[io];
        }
        // ---- B: d_x on rows of the tile (x' in [x0-4, x0+68)), d_y on columns (y' in [y0-4, y0+BY+4))
        for (int k = tid; k < BY * BW; k += BX * BY) {
            const int r = k / BW, j = k - r * BW;
            const int gx = x0 - 4 + j, gy = y0 + r;
            float v1 = 0.0f, v2 = 0.0f;
            if (interior(g, gx, gy, p)) {
                float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    a1 += c * (T1[cur][r + 8][j + 3 + n] - T1[cur][r + 8][j + 4 - n]);
                    a2 += c * (T2[cur][r + 8][j + 3 + n] - T2[cur][r + 8][j + 4 - n]);
                }
                const float bb = Bt[cur][r + 4][j];
                v1 = a1 * invdx * bb;
                v2 = a2 * invdx * bb;
            }
            X1[r][j] = v1;
            X2[r][j] = v2;
        }
        for (int k = tid; k < BH * BX; k += BX * BY) {
            const int j = k / BX, c0 = k - j * BX;
            const int gx = x0 + c0, gy = y0 - 4 + j;
            float v1 = 0.0f, v2 = 0.0f;
            if (interior(g, gx, gy, p)) {
                float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    a1 += c * (T1[cur][j + 3 + n][c0 + 8] - T1[cur][j + 4 - n][c0 + 8]);
                    a2 += c * (T2[cur][j + 3 + n][c0 + 8] - T2[cur][j + 4 - n][c0 + 8]);
                }
                const float bb = Bt[cur][j][c0 + 4];
                v1 = a1 * invdy * bb;
                v2 = a2 * invdy * bb;
            }
            Y1[j][c0] = v1;
            Y2[j][c0] = v2;
        }
        __syncthreads();
        // ---- C: plane-p contributions
        if (active) {
            float fx1 = 0, fx2 = 0, fy1 = 0, fy2 = 0, fxy1 = 0, fxy2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                fx1 += c * (X1[ty][tx + 4 + n] - X1[ty][tx + 5 - n]);
                fx2 += c * (X2[ty][tx + 4 + n] - X2[ty][tx + 5 - n]);
                fy1 += c * (Y1[ty + 4 + n][tx] - Y1[ty + 5 - n][tx]);
                fy2 += c * (Y2[ty + 4 + n][tx] - Y2[ty + 5 - n][tx]);
                const float e = ecen(n);
                fxy1 += e * ((T1[cur][ty + 8 + n][tx + 8 + n] - T1[cur][ty + 8 - n][tx + 8 + n]) +
                             (T1[cur][ty + 8 - n][tx + 8 - n] - T1[cur][ty + 8 + n][tx + 8 - n]));
                fxy2 += e * ((T2[cur][ty + 8 + n][tx + 8 + n] - T2[cur][ty + 8 - n][tx + 8 + n]) +
                             (T2[cur][ty + 8 - n][tx + 8 - n] - T2[cur][ty + 8 + n][tx + 8 - n]));
                // plane p at x-offset +n feeds output p - n (q[4 - n]); at -n output p + n (q[4 + n]), sign flipped
                const float dx1 = T1[cur][ty + 8][tx + 8 + n] - T1[cur][ty + 8][tx + 8 - n];
                const float dx2 = T2[cur][ty + 8][tx + 8 + n] - T2[cur][ty + 8][tx + 8 - n];
                const float dy1 = T1[cur][ty + 8 + n][tx + 8] - T1[cur][ty + 8 - n][tx + 8];
                const float dy2 = T2[cur][ty + 8 + n][tx + 8] - T2[cur][ty + 8 - n][tx + 8];
                qx1[4 - n] += e * dx1;
                qx1[4 + n] -= e * dx1;
                qx2[4 - n] += e * dx2;
                qx2[4 + n] -= e * dx2;
                qy1[4 - n] += e * dy1;
                qy1[4 + n] -= e * dy1;
                qy2[4 - n] += e * dy2;
                qy2[4 + n] -= e * dy2;
            }
            xx1[4] = fx1 * invdx;
            xx2[4] = fx2 * invdx;
            yy1[4] = fy1 * invdy;
            yy2[4] = fy2 * invdy;
            xy1[4] = fxy1 * i4xy;
            xy2[4] = fxy2 * i4xy;
        }
        // ---- D: d_z(o+4) and the finalisation of plane o
        {
            float v1 = 0.0f, v2 = 0.0f;
            if (dzok) {
                float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    a1 += c * (sq1[R + n - 1] - sq1[R - n]);
                    a2 += c * (sq2[R + n - 1] - sq2[R - n]);
                }
                v1 = a1 * invdz * bd;
                v2 = a2 * invdz * bd;
            }
            dz1[7] = v1;
            dz2[7] = v2;
        }
        float unew = 0.0f, wnew = 0.0f;
        if (fin) {
            float fz1 = 0.0f, fz2 = 0.0f;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                fz1 += c * (dz1[3 + n] - dz1[4 - n]);
                fz2 += c * (dz2[3 + n] - dz2[4 - n]);
            }
            const float rho = 1.0f / bo;
            const float f1xx = xx1[0] * rho, f2xx = xx2[0] * rho;
            const float f1yy = yy1[0] * rho, f2yy = yy2[0] * rho;
            const float f1zz = fz1 * invdz * rho, f2zz = fz2 * invdz * rho;
            const float f1xy = xy1[0], f2xy = xy2[0];
            const float f1zx = qx1[0] * i4zx, f2zx = qx2[0] * i4zx;
            const float f1yz = qy1[0] * i4yz, f2yz = qy2[0] * i4yz;
            const float rx2 = vx * vx, ry2 = vy * vy, rz2 = vz * vz;
            const float rxy = 2.0f * vx * vy, ryz = 2.0f * vy * vz, rzx = 2.0f * vz * vx;
            const float V1 = rx2 * f1xx + ry2 * f1yy + rz2 * f1zz + rxy * f1xy + rzx * f1zx + ryz * f1yz;
            const float H1 = f1xx + f1yy + f1zz - V1;
            const float V2 = rx2 * f2xx + ry2 * f2yy + rz2 * f2zz + rxy * f2xy + rzx * f2zx + ryz * f2yz;
            const float H2 = f2xx + f2yy + f2zz - V2;
            const float r3 = mm1 * H1 + mm4 * V1 + mm2 * V2;
            const float r4 = mm2 * H1 + mm3 * V2 + mm4 * H2;
            unew = uo + r3 * dt;
            wnew = wo + r4 * dt;
            if (io == src_idx) {
                unew += src_val;
                wnew += src_val;
            }
            u[io] = unew;
            w[io] = wnew;
            float a = 1.0f, bb = 1.0f;
            if (sponge) {
                a = fminf(axy, abz[o].a);
                bb = fminf(bxy, abz[o].b);
            }
            s1o[io] = a * sq1[0] + bb * unew * dt;
            s2o[io] = a * sq2[0] + bb * wnew * dt;
        }
        if (snap.mode != TTI_SNAP_NONE && emit && iy <= g.iyend) {
            const long row = (long)(o - g.izbeg) * snap.ldy + (iy - g.iybeg);
            const long idx = row * snap.ldx + (ix - g.ixbeg);
            const long sidx = row * (snap.ldx / BX) + (x0 - g.ixbeg) / BX;
            snap_store(snap.mode, snap.a, snap.sa, idx, sidx, active, unew);
            snap_store(snap.mode, snap.b, snap.sb, idx, sidx, active, wnew);
        }
        // ---- E: next tile into the other buffer, prefetch the one after
        if (p + 1 <= z1 + R - 1)
            commit(cur ^ 1);
        if (p + 2 <= z1 + R - 1)
            fetch(p + 2);
        __syncthreads();
        cur ^= 1;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            sq1[k] = sq1[k + 1];
            sq2[k] = sq2[k + 1];
            dz1[k] = dz1[k + 1];
            dz2[k] = dz2[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            xx1[k] = xx1[k + 1];
            xx2[k] = xx2[k + 1];
            yy1[k] = yy1[k + 1];
            yy2[k] = yy2[k + 1];
            xy1[k] = xy1[k + 1];
            xy2[k] = xy2[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            qx1[k] = qx1[k + 1];
            qx2[k] = qx2[k + 1];
            qy1[k] = qy1[k + 1];
            qy2[k] = qy2[k + 1];
        }
        qx1[8] = qx2[8] = qy1[8] = qy2[8] = 0.0f;
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Adjoint step (reference order, kernels/kernel_type2_grad_snap.cu:242-427 + kernels/kernel_type2_adj.cu:100-260 +
// kernels/kernel_base.cu:82-104):
//   r  = sum_a D+_a (b D-_a S_a) + sum_{a<b} X_ab M_ab,   S_a = rho (lap + r_a^2 hz), M_ab = 2 r_a r_b hz
//   u' = u + dt r            (sponge: u' = A u + B dt r)
//   sigma' = A sigma + B dt u'
//   grad += u' u_snap + w' w_snap   (FP64, when a snapshot is given)
// Same pipeline as the forward kernel; the tile stage holds the pointwise temporaries S_x, S_y, M_xy, M_zx, M_yz
// and b of plane p on a plus-shaped footprint, computed from the 10 prefetched input quads when they are stored.
// ------------------------------------------------------------------------------------------------------------------
struct Temps
{
    float lap1, hz1, lap2, hz2;
};

__device__ __forceinline__ Temps
temps(float f1, float f2, float mm1, float mm2, float mm3, float mm4)
{
    Temps t;
    t.lap1 = mm1 * f1 + mm2 * f2;
    t.hz1 = (-mm1 + mm4) * f1 - mm2 * f2;
    t.lap2 = mm4 * f2;
    t.hz2 = mm2 * f1 + (-mm4 + mm3) * f2;
    return t;
}

struct __align__(16) AdjTiles
{
    float SX1[BY][TW], SX2[BY][TW];   // S_x at (x0-8+lx, y0+r)
    float SY1[TH][BX], SY2[TH][BX];   // S_y at (x0+c, y0-8+ly)
    float MXY1[BH][BW], MXY2[BH][BW]; // M_xy at (x0-4+lx, y0-4+ly)
    float MZX1[BY][BW], MZX2[BY][BW]; // M_zx at (x0-4+lx, y0+r)
    float MYZ1[BH][BX], MYZ2[BH][BX]; // M_yz at (x0+c, y0-4+ly)
    float Bt[BH][BW];                 // b at (x0-4+lx, y0-4+ly)
};

__global__ __launch_bounds__(BX* BY) void k_fused_adj(
    const float* __restrict__ a1, const float* __restrict__ a2, float* __restrict__ a1o, float* __restrict__ a2o,
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ m1, const float* __restrict__ m2,
    const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ vrx,
    const float* __restrict__ vry, const float* __restrict__ vrz, const float* __restrict__ b,
    const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, Box g, float dt,
    float invdx, float invdy, float invdz, float i4xy, float i4yz, float i4zx, Snap snap, double* __restrict__ grad,
    int bz, int ntx, int nty, int ntz)
{
    __shared__ AdjTiles S[2];
    __shared__ float X1[BY][BW], X2[BY][BW]; // d_x
    __shared__ float Y1[BH][BX], Y2[BH][BX]; // d_y

    int bxi, byi, bzi;
    if (!decode_tile(ntx, nty, ntz, bxi, byi, bzi))
        return;
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    const int x0 = g.ixbeg + bxi * BX, y0 = g.iybeg + byi * BY, z0 = g.izbeg + bzi * bz;
    const int z1 = min(z0 + bz, g.izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= g.ixend && iy <= g.iyend;
    const long sz = g.ldx * g.ldy;
    const long col = (long)iy * g.ldx + ix;
    const bool sponge = abx != nullptr;
    float axy = 1.0f, bxy = 1.0f;
    if (sponge && active) {
        axy = fminf(abx[ix].a, aby[iy].a);
        bxy = fminf(abx[ix].b, aby[iy].b);
    }
    const long gstride_y = (long)(g.ixend - g.ixbeg + 1), gstride_z = gstride_y * (g.iyend - g.iybeg + 1);

    // quad map of the (TW x TH) footprint; quads are 4-aligned in lx, as are the footprint regions
    const bool tq_on = tid < TQ;
    const int tq_row = tid / TQ_PER_ROW, tq_lx = 4 * (tid - tq_row * TQ_PER_ROW);
    const bool row8 = tq_row >= 8 && tq_row < 8 + BY;
    const bool col8 = tq_lx >= 8 && tq_lx < 8 + BX;
    const bool sq4 = tq_lx >= 4 && tq_lx < 4 + BW && tq_row >= 4 && tq_row < 4 + BH;
    const bool tq_need = tq_on && (row8 || col8 || sq4);
    const int tq_gx = x0 - 8 + tq_lx, tq_gy = y0 - 8 + tq_row;
    const bool tq_ok = tq_need && tq_gy >= g.iybeg - R && tq_gy <= g.iyend + R;
    const long tq_off = (long)tq_gy * g.ldx + tq_gx;
    float4 qa1, qa2, qm1, qm2, qm3, qm4, qrx, qry, qrz, qb;
    auto fetch = [&](int pp) {
        const long i = (long)pp * sz + tq_off;
        qa1 = load_quad(a1, i, tq_gx, g, tq_ok);
        qa2 = load_quad(a2, i, tq_gx, g, tq_ok);
        qm1 = load_quad(m1, i, tq_gx, g, tq_ok);
        qm2 = load_quad(m2, i, tq_gx, g, tq_ok);
        qm3 = load_quad(m3, i, tq_gx, g, tq_ok);
        qm4 = load_quad(m4, i, tq_gx, g, tq_ok);
        qrx = load_quad(vrx, i, tq_gx, g, tq_ok);
        qry = load_quad(vry, i, tq_gx, g, tq_ok);
        qrz = load_quad(vrz, i, tq_gx, g, tq_ok);
        qb = load_quad(b, i, tq_gx, g, tq_ok);
    };
    auto commit = [&](int buf) {
        if (!tq_need)
            return;
        AdjTiles& t = S[buf];
        const float* A1 = &qa1.x;
        const float* A2 = &qa2.x;
        const float* M1 = &qm1.x;
        const float* M2 = &qm2.x;
        const float* M3 = &qm3.x;
        const float* M4 = &qm4.x;
        const float* RX = &qrx.x;
        const float* RY = &qry.x;
        const float* RZ = &qrz.x;
        const float* BB = &qb.x;
        float sx1[4], sx2[4], sy1[4], sy2[4], mxy1[4], mxy2[4], mzx1[4], mzx2[4], myz1[4], myz2[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const Temps tt = temps(A1[e], A2[e], M1[e], M2[e], M3[e], M4[e]);
            const float rho = BB[e] != 0.0f ? 1.0f / BB[e] : 0.0f; // zero outside the allocation
            const float vx = RX[e], vy = RY[e], vz = RZ[e];
            sx1[e] = rho * (tt.lap1 + vx * vx * tt.hz1);
            sx2[e] = rho * (tt.lap2 + vx * vx * tt.hz2);
            sy1[e] = rho * (tt.lap1 + vy * vy * tt.hz1);
            sy2[e] = rho * (tt.lap2 + vy * vy * tt.hz2);
            mxy1[e] = 2.0f * vx * vy * tt.hz1;
            mxy2[e] = 2.0f * vx * vy * tt.hz2;
            mzx1[e] = 2.0f * vz * vx * tt.hz1;
            mzx2[e] = 2.0f * vz * vx * tt.hz2;
            myz1[e] = 2.0f * vy * vz * tt.hz1;
            myz2[e] = 2.0f * vy * vz * tt.hz2;
        }
        auto st = [](float* dst, const float* v) {
            *reinterpret_cast<float4*>(dst) = make_float4(v[0], v[1], v[2], v[3]);
        };
        if (row8) {
            st(&t.SX1[tq_row - 8][tq_lx], sx1);
            st(&t.SX2[tq_row - 8][tq_lx], sx2);
        }
        if (col8) {
            st(&t.SY1[tq_row][tq_lx - 8], sy1);
            st(&t.SY2[tq_row][tq_lx - 8], sy2);
        }
        if (sq4) {
            st(&t.MXY1[tq_row - 4][tq_lx - 4], mxy1);
            st(&t.MXY2[tq_row - 4][tq_lx - 4], mxy2);
            st(&t.Bt[tq_row - 4][tq_lx - 4], BB);
            if (row8) {
                st(&t.MZX1[tq_row - 8][tq_lx - 4], mzx1);
                st(&t.MZX2[tq_row - 8][tq_lx - 4], mzx2);
            }
            if (col8) {
                st(&t.MYZ1[tq_row - 4][tq_lx - 8], myz1);
                st(&t.MYZ2[tq_row - 4][tq_lx - 8], myz2);
            }
        }
    };

    float sz1[8], sz2[8];                 // S_z(o + k)
    float dz1[8], dz2[8];                 // d_z(o - 3 + k)
    float fx1[5], fx2[5], fy1[5], fy2[5]; // D+_x d_x, D+_y d_y at plane o + k
    float xy1[5], xy2[5];                 // X_xy M_xy
    float q1[9], q2[9];                   // X_zx + X_yz partial sums
#pragma unroll
    for (int k = 0; k < 8; ++k)
        dz1[k] = dz2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        fx1[k] = fx2[k] = fy1[k] = fy2[k] = xy1[k] = xy2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; ++k)
        q1[k] = q2[k] = 0.0f;

    auto col_sz = [&](int q, float& v1, float& v2) {
        v1 = v2 = 0.0f;
        if (active && q >= g.izbeg - R && q <= g.izend + R) {
            const long i = (long)q * sz + col;
            const Temps t = temps(a1[i], a2[i], m1[i], m2[i], m3[i], m4[i]);
            const float rho = 1.0f / b[i];
            const float vz = vrz[i];
            v1 = rho * (t.lap1 + vz * vz * t.hz1);
            v2 = rho * (t.lap2 + vz * vz * t.hz2);
        }
    };
#pragma unroll
    for (int k = 0; k < 7; ++k)
        col_sz(z0 - 8 + k, sz1[k], sz2[k]);
    int cur = 0;
    fetch(z0 - R);
    commit(0);
    fetch(z0 - R + 1);
    __syncthreads();

    for (int o = z0 - 8; o < z1; ++o) {
        const int p = o + R;
        const bool fin = o >= z0 && active;
        // ---- A: column and pointwise loads
        col_sz(o + 7, sz1[7], sz2[7]);
        const int qd = o + 4;
        const bool dzok = active && qd >= g.izbeg && qd <= g.izend;
        const float bd = dzok ? b[(long)qd * sz + col] : 0.0f;
        const long io = (long)o * sz + col;
        float uo = 0, wo = 0, ao1 = 0, ao2 = 0, su = 0, sw = 0;
        const bool img = fin && snap.mode != TTI_SNAP_NONE && grad != nullptr;
        long gi = 0;
        double g0 = 0.0;
        if (fin) {
            uo = u[io];
            wo = w[io];
            ao1 = a1[io];
            ao2 = a2[io];
        }
        if (img) {
            const long row = (long)(o - g.izbeg) * snap.ldy + (iy - g.iybeg);
            const long idx = row * snap.ldx + (ix - g.ixbeg);
            const long sidx = row * (snap.ldx / BX) + (x0 - g.ixbeg) / BX;
            su = snap_load(snap.mode, snap.a, snap.sa, idx, sidx);
            sw = snap_load(snap.mode, snap.b, snap.sb, idx, sidx);
            gi = (long)(o - g.izbeg) * gstride_z + (long)(iy - g.iybeg) * gstride_y + (ix - g.ixbeg);
            g0 = grad[gi];
        }
        const AdjTiles& t = S[cur];
        // ---- B: d_x, d_y of plane p
        for (int k = tid; k < BY * BW; k += BX * BY) {
            const int r = k / BW, j = k - r * BW;
            const int gx = x0 - 4 + j, gy = y0 + r;
            float v1 = 0.0f, v2 = 0.0f;
            if (interior(g, gx, gy, p)) {
                float s_1 = 0.0f, s_2 = 0.0f;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    s_1 += c * (t.SX1[r][j + 3 + n] - t.SX1[r][j + 4 - n]);
                    s_2 += c * (t.SX2[r][j + 3 + n] - t.SX2[r][j + 4 - n]);
                }
                const float bb = t.Bt[r + 4][j];
                v1 = s_1 * invdx * bb;
                v2 = s_2 * invdx * bb;
            }
            X1[r][j] = v1;
            X2[r][j] = v2;
        }
        for (int k = tid; k < BH * BX; k += BX * BY) {
            const int j = k / BX, c0 = k - j * BX;
            const int gx = x0 + c0, gy = y0 - 4 + j;
            float v1 = 0.0f, v2 = 0.0f;
            if (interior(g, gx, gy, p)) {
                float s_1 = 0.0f, s_2 = 0.0f;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    s_1 += c * (t.SY1[j + 3 + n][c0] - t.SY1[j + 4 - n][c0]);
                    s_2 += c * (t.SY2[j + 3 + n][c0] - t.SY2[j + 4 - n][c0]);
                }
                const float bb = t.Bt[j][c0 + 4];
                v1 = s_1 * invdy * bb;
                v2 = s_2 * invdy * bb;
            }
            Y1[j][c0] = v1;
            Y2[j][c0] = v2;
        }
        __syncthreads();
        // ---- C: plane-p contributions
        if (active) {
            float gx1 = 0, gx2 = 0, gy1 = 0, gy2 = 0, hxy1 = 0, hxy2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                gx1 += c * (X1[ty][tx + 4 + n] - X1[ty][tx + 5 - n]);
                gx2 += c * (X2[ty][tx + 4 + n] - X2[ty][tx + 5 - n]);
                gy1 += c * (Y1[ty + 4 + n][tx] - Y1[ty + 5 - n][tx]);
                gy2 += c * (Y2[ty + 4 + n][tx] - Y2[ty + 5 - n][tx]);
                const float e = ecen(n);
                hxy1 += e * (t.MXY1[ty + 4 + n][tx + 4 + n] - t.MXY1[ty + 4 + n][tx + 4 - n] +
                             t.MXY1[ty + 4 - n][tx + 4 - n] - t.MXY1[ty + 4 - n][tx + 4 + n]);
                hxy2 += e * (t.MXY2[ty + 4 + n][tx + 4 + n] - t.MXY2[ty + 4 + n][tx + 4 - n] +
                             t.MXY2[ty + 4 - n][tx + 4 - n] - t.MXY2[ty + 4 - n][tx + 4 + n]);
                const float d1 = i4zx * (t.MZX1[ty][tx + 4 + n] - t.MZX1[ty][tx + 4 - n]) +
                                 i4yz * (t.MYZ1[ty + 4 + n][tx] - t.MYZ1[ty + 4 - n][tx]);
                const float d2 = i4zx * (t.MZX2[ty][tx + 4 + n] - t.MZX2[ty][tx + 4 - n]) +
                                 i4yz * (t.MYZ2[ty + 4 + n][tx] - t.MYZ2[ty + 4 - n][tx]);
                q1[4 - n] += e * d1;
                q1[4 + n] -= e * d1;
                q2[4 - n] += e * d2;
                q2[4 + n] -= e * d2;
            }
            fx1[4] = gx1 * invdx;
            fx2[4] = gx2 * invdx;
            fy1[4] = gy1 * invdy;
            fy2[4] = gy2 * invdy;
            xy1[4] = hxy1 * i4xy;
            xy2[4] = hxy2 * i4xy;
        }
        // ---- D: d_z(o+4), finalise plane o
        {
            float v1 = 0.0f, v2 = 0.0f;
            if (dzok) {
                float s_1 = 0.0f, s_2 = 0.0f;
#pragma unroll
                for (int n = 1; n <= R; ++n) {
                    const float c = cstg(n);
                    s_1 += c * (sz1[R + n - 1] - sz1[R - n]);
                    s_2 += c * (sz2[R + n - 1] - sz2[R - n]);
                }
                v1 = s_1 * invdz * bd;
                v2 = s_2 * invdz * bd;
            }
            dz1[7] = v1;
            dz2[7] = v2;
        }
        if (fin) {
            float gz1 = 0.0f, gz2 = 0.0f;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                gz1 += c * (dz1[3 + n] - dz1[4 - n]);
                gz2 += c * (dz2[3 + n] - dz2[4 - n]);
            }
            const float r3 = fx1[0] + fy1[0] + gz1 * invdz + xy1[0] + q1[0];
            const float r4 = fx2[0] + fy2[0] + gz2 * invdz + xy2[0] + q2[0];
            float a = 1.0f, bb = 1.0f;
            if (sponge) {
                a = fminf(axy, abz[o].a);
                bb = fminf(bxy, abz[o].b);
            }
            const float un = sponge ? a * uo + bb * dt * r3 : uo + r3 * dt;
            const float wn = sponge ? a * wo + bb * dt * r4 : wo + r4 * dt;
            u[io] = un;
            w[io] = wn;
            a1o[io] = a * ao1 + bb * un * dt;
            a2o[io] = a * ao2 + bb * wn * dt;
            if (img)
                grad[gi] = g0 + ((double)un * (double)su + (double)wn * (double)sw);
        }
        // ---- E: temporaries of plane p+1 into the other buffer, prefetch plane p+2
        if (p + 1 <= z1 + R - 1)
            commit(cur ^ 1);
        if (p + 2 <= z1 + R - 1)
            fetch(p + 2);
        __syncthreads();
        cur ^= 1;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            sz1[k] = sz1[k + 1];
            sz2[k] = sz2[k + 1];
            dz1[k] = dz1[k + 1];
            dz2[k] = dz2[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            fx1[k] = fx1[k + 1];
            fx2[k] = fx2[k + 1];
            fy1[k] = fy1[k + 1];
            fy2[k] = fy2[k + 1];
            xy1[k] = xy1[k + 1];
            xy2[k] = xy2[k + 1];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            q1[k] = q1[k + 1];
            q2[k] = q2[k + 1];
        }
        q1[8] = q2[8] = 0.0f;
    }
}

// ------------------------------------------------------------------------------------------------------------------
// Receivers and sources
// ------------------------------------------------------------------------------------------------------------------

// out[k] = -(2 s1 + s2)/3 at idx[k] (pressure recorded at receivers, simulator2/simulator2SetUp.cu:31-42)
__global__ __launch_bounds__(256) void k_gather_pressure(const float* __restrict__ s1, const float* __restrict__ s2,
                                                         const long* __restrict__ idx, int n, float* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        const long i = idx[k];
        out[k] = -(2.0f * s1[i] + s2[i]) / 3.0f;
    }
}

// Transpose of the pressure recording, applied after an adjoint step: u += -2/3 r, w += -1/3 r at receiver k and the
// already-updated sigma_out gets B dt times the same increment (the adjoint step's sigma' = A sigma + B dt u').
// Receivers must be distinct points. Optionally adds the increment's share of the imaging condition.
__global__ __launch_bounds__(256) void k_inject_residual(float* __restrict__ u, float* __restrict__ w,
                                                         float* __restrict__ a1o, float* __restrict__ a2o,
                                                         const long* __restrict__ idx, const int* __restrict__ ijk,
                                                         int n, const float* __restrict__ res, float dt,
                                                         const ab2* __restrict__ abx, const ab2* __restrict__ aby,
                                                         const ab2* __restrict__ abz, Box g, Snap snap,
                                                         double* __restrict__ grad)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const long i = idx[k];
    const int x = ijk[3 * k], y = ijk[3 * k + 1], z = ijk[3 * k + 2];
    const float du = -2.0f / 3.0f * res[k], dw = -1.0f / 3.0f * res[k];
    float bb = 1.0f;
    if (abx != nullptr)
        bb = fminf(fminf(abx[x].b, aby[y].b), abz[z].b);
    u[i] += du;
    w[i] += dw;
    a1o[i] += bb * du * dt;
    a2o[i] += bb * dw * dt;
    if (snap.mode != TTI_SNAP_NONE && grad != nullptr) {
        const long row = (long)(z - g.izbeg) * snap.ldy + (y - g.iybeg);
        const long sidx = row * (snap.ldx / BX) + (x - g.ixbeg) / BX;
        const long si = row * snap.ldx + (x - g.ixbeg);
        const float su = snap_load(snap.mode, snap.a, snap.sa, si, sidx);
        const float sw = snap_load(snap.mode, snap.b, snap.sb, si, sidx);
        const long gstride_y = (long)(g.ixend - g.ixbeg + 1), gstride_z = gstride_y * (g.iyend - g.iybeg + 1);
        const long gi = (long)(z - g.izbeg) * gstride_z + (long)(y - g.iybeg) * gstride_y + (x - g.ixbeg);
        grad[gi] += (double)du * (double)su + (double)dw * (double)sw;
    }
}

// Decode a snapshot field back to FP32 (interior layout nz x ny x nx) for inspection and tests.
__global__ __launch_bounds__(256) void k_snap_decode(int mode, const void* __restrict__ arr,
                                                     const float* __restrict__ scales, long sldx, long sldy, int nx,
                                                     int ny, int nz, float* __restrict__ out)
{
    const long n = (long)nx * ny * nz;
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const int x = (int)(k % nx);
    const long r = k / nx;
    const int y = (int)(r % ny), z = (int)(r / ny);
    const long row = (long)z * sldy + y;
    out[k] = snap_load(mode, arr, scales, row * sldx + x, row * (sldx / BX) + x / BX);
}

} // namespace fused
} // namespace tti

// ======================================================================================================================
// C ABI (include/tti_fused.h)
// ======================================================================================================================
using namespace tti;
using namespace tti::fused;

namespace {

// The fused kernels stage tiles with 16-byte loads: the interior must start on a 4-float boundary, rows must be a
// multiple of 4 floats and every volume 16-byte aligned.
bool
box_ok(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy)
{
    return ixbeg >= R && iybeg >= R && izbeg >= R && ixend >= ixbeg && iyend >= iybeg && izend >= izbeg &&
           ixend + R < ldimx && iyend + R < ldimy && ixbeg % 4 == 0 && ldimx % 4 == 0;
}

bool
aligned16(std::initializer_list<const void*> ps)
{
    for (const void* p : ps)
        if (reinterpret_cast<uintptr_t>(p) % 16 != 0)
            return false;
    return true;
}

bool
snap_ok(const tti_snapshot* s, int nx, int ny)
{
    if (s == nullptr || s->mode == TTI_SNAP_NONE)
        return true;
    if (s->mode < TTI_SNAP_FP32 || s->mode > TTI_SNAP_BF16 || !s->field3 || !s->field4)
        return false;
    if (s->ldimx % 64 != 0 || s->ldimx < nx || s->ldimy < ny)
        return false;
    if (s->mode == TTI_SNAP_FP16 && (!s->scale3 || !s->scale4))
        return false;
    return true;
}

Snap
to_snap(const tti_snapshot* s)
{
    Snap r{ TTI_SNAP_NONE, nullptr, nullptr, nullptr, nullptr, 64, 1 };
    if (s && s->mode != TTI_SNAP_NONE) {
        r.mode = s->mode;
        r.a = s->field3;
        r.b = s->field4;
        r.sa = s->scale3;
        r.sb = s->scale4;
        r.ldx = s->ldimx;
        r.ldy = s->ldimy;
    }
    return r;
}

struct Launch
{
    dim3 grid;
    int ntx, nty, ntz, bz;
};

Launch
plan(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int bz)
{
    Launch L;
    const int nx = ixend - ixbeg + 1, ny = iyend - iybeg + 1, nz = izend - izbeg + 1;
    L.bz = bz > 0 ? bz : 128;
    L.ntx = (nx + BX - 1) / BX;
    L.nty = (ny + BY - 1) / BY;
    L.ntz = (nz + L.bz - 1) / L.bz;
    const int nt = L.ntx * L.nty * L.ntz;
    L.grid = dim3(((nt + 7) / 8) * 8); // multiple of 8 so the XCD round-robin decode covers every tile
    return L;
}

} // namespace

extern "C" {

int
tti_fused_fwd_step(const float* field1, const float* field2, float* field1_out, float* field2_out, float* field3,
                   float* field4, const float* model1, const float* model2, const float* model3, const float* model4,
                   const float* rx, const float* ry, const float* rz, const float* Model11, const void* ABx,
                   const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                   int ldimx, int ldimy, float dt, double dx, double dy, double dz, long src_index, float src_value,
                   const tti_snapshot* snap, int zblock, void* stream)
{
    if (!field1 || !field2 || !field1_out || !field2_out || !field3 || !field4 || !model1 || !model2 || !model3 ||
        !model4 || !rx || !ry || !rz || !Model11 || field1 == field1_out || field2 == field2_out)
        return TTI_EINVAL;
    if (!box_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !snap_ok(snap, ixend - ixbeg + 1, iyend - iybeg + 1))
        return TTI_EINVAL;
    if ((ABx || ABy || ABz) && !(ABx && ABy && ABz))
        return TTI_EINVAL;
    if (!aligned16({ field1, field2, Model11 }))
        return TTI_EINVAL;
    const Box g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy };
    const Launch L = plan(ixbeg, ixend, iybeg, iyend, izbeg, izend, zblock);
    k_fused_fwd<<<L.grid, dim3(BX, BY), 0, (hipStream_t)stream>>>(
        field1, field2, field1_out, field2_out, field3, field4, model1, model2, model3, model4, rx, ry, rz, Model11,
        (const ab2*)ABx, (const ab2*)ABy, (const ab2*)ABz, g, dt, (float)(1.0 / dx), (float)(1.0 / dy),
        (float)(1.0 / dz), (float)(1.0 / (4 * dx * dy)), (float)(1.0 / (4 * dy * dz)), (float)(1.0 / (4 * dx * dz)),
        src_index, src_value, to_snap(snap), L.bz, L.ntx, L.nty, L.ntz);
    return launch_status();
}

int
tti_fused_adj_step(const float* field1, const float* field2, float* field1_out, float* field2_out, float* field3,
                   float* field4, const float* model1, const float* model2, const float* model3, const float* model4,
                   const float* rx, const float* ry, const float* rz, const float* Model11, const void* ABx,
                   const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                   int ldimx, int ldimy, float dt, double dx, double dy, double dz, const tti_snapshot* snap,
                   double* grad, int zblock, void* stream)
{
    if (!field1 || !field2 || !field1_out || !field2_out || !field3 || !field4 || !model1 || !model2 || !model3 ||
        !model4 || !rx || !ry || !rz || !Model11 || field1 == field1_out || field2 == field2_out)
        return TTI_EINVAL;
    if (!box_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !snap_ok(snap, ixend - ixbeg + 1, iyend - iybeg + 1))
        return TTI_EINVAL;
    if ((ABx || ABy || ABz) && !(ABx && ABy && ABz))
        return TTI_EINVAL;
    if (!aligned16({ field1, field2, model1, model2, model3, model4, rx, ry, rz, Model11 }))
        return TTI_EINVAL;
    const Box g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy };
    const Launch L = plan(ixbeg, ixend, iybeg, iyend, izbeg, izend, zblock);
    k_fused_adj<<<L.grid, dim3(BX, BY), 0, (hipStream_t)stream>>>(
        field1, field2, field1_out, field2_out, field3, field4, model1, model2, model3, model4, rx, ry, rz, Model11,
        (const ab2*)ABx, (const ab2*)ABy, (const ab2*)ABz, g, dt, (float)(1.0 / dx), (float)(1.0 / dy),
        (float)(1.0 / dz), (float)(1.0 / (4 * dx * dy)), (float)(1.0 / (4 * dy * dz)), (float)(1.0 / (4 * dx * dz)),
        to_snap(snap), grad, L.bz, L.ntx, L.nty, L.ntz);
    return launch_status();
}

int
tti_gather_pressure(const float* field1, const float* field2, const long* index, int n, float* out, void* stream)
{
    if (n < 0 || (n > 0 && (!field1 || !field2 || !index || !out)))
        return TTI_EINVAL;
    if (n == 0)
        return TTI_OK;
    k_gather_pressure<<<dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream>>>(field1, field2, index, n, out);
    return launch_status();
}

int
tti_inject_residual(float* field3, float* field4, float* field1_out, float* field2_out, const long* index,
                    const int* ijk, int n, const float* residual, float dt, const void* ABx, const void* ABy,
                    const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                    const tti_snapshot* snap, double* grad, void* stream)
{
    if (n < 0 || (n > 0 && (!field3 || !field4 || !field1_out || !field2_out || !index || !ijk || !residual)))
        return TTI_EINVAL;
    if ((ABx || ABy || ABz) && !(ABx && ABy && ABz))
        return TTI_EINVAL;
    if (!snap_ok(snap, ixend - ixbeg + 1, iyend - iybeg + 1))
        return TTI_EINVAL;
    if (n == 0)
        return TTI_OK;
    const Box g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, 0, 0 };
    k_inject_residual<<<dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream>>>(
        field3, field4, field1_out, field2_out, index, ijk, n, residual, dt, (const ab2*)ABx, (const ab2*)ABy,
        (const ab2*)ABz, g, to_snap(snap), grad);
    return launch_status();
}

int
tti_snapshot_decode(const tti_snapshot* snap, int which, int nx, int ny, int nz, float* out, void* stream)
{
    if (!snap || snap->mode == TTI_SNAP_NONE || !snap_ok(snap, nx, ny) || !out || nx < 1 || ny < 1 || nz < 1 ||
        (which != 3 && which != 4))
        return TTI_EINVAL;
    const long n = (long)nx * ny * nz;
    const void* arr = which == 3 ? snap->field3 : snap->field4;
    const float* sc = which == 3 ? snap->scale3 : snap->scale4;
    k_snap_decode<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream>>>(
        snap->mode, arr, sc, (long)snap->ldimx, (long)snap->ldimy, nx, ny, nz, out);
    return launch_status();
}

} // extern "C"
