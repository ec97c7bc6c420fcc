#pragma once
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
#include <type_traits>

namespace tti {

namespace fused {

constexpr int BX = 64;
constexpr int BY = 8;
constexpr int TW = BX + 16, TH = BY + 16; // sigma tile with halo 8
constexpr int BW = BX + 8, BH = BY + 8;   // halo-4 tile
constexpr int SNAP_SEG = 64;              // points per x-row segment of an FP16 snapshot (one scale each)

struct Snap
{
    int mode; // TTI_SNAP_NONE / FP32 / FP16 / BF16
    void* a;  // u
    void* b;  // w
    float* sa;
    float* sb;
    long ldx; // row length of the snapshot arrays (multiple of 64)
    long ldy;
    void* c;  // sigma_H of the step (density image), or null
    void* d;  // sigma_V
    float* sc;
    float* sd;
};

struct Box
{
    int ixbeg, ixend, iybeg, iyend, izbeg, izend;
    long ldx, ldy;
    int fs; // stress-free surface at plane izbeg (the z derivatives of planes <= izbeg mirrored, see mirror_dz)
};

// Free-surface mirror of the z-derivative queue (kernels/kernel_type2_small.cu:5-40): d_z is even about
// izbeg + 1/2, d_z(iz) = d_z(2 izbeg + 1 - iz) for every iz <= izbeg. When the queue window q[c .. c+7] holds
// d_z(izbeg - 3 .. izbeg + 4) (output plane o = izbeg), its first four slots take the last four in reverse; the
// outputs after it see the mirrored values as the window moves on.
template <int C, int N>
__device__ __forceinline__ void
mirror_dz(float (&q)[N])
{
#pragma unroll
    for (int k = 0; k < 4; ++k)
        q[C + k] = q[C + 7 - k];
}

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

// Block schedule within one z chunk: tile r -> (bxi, byi). strip <= 0 or >= ntx: x fastest over full rows; otherwise
// the chunk is cut into vertical strips of `strip` x tiles (the last one narrower), x fastest within a strip, then y,
// strip after strip. Narrow strips keep the blocks that run at the same time on one XCD closer together in x-y
// (their halos overlap in that XCD's L2); the autotuner picks the width.
__device__ __forceinline__ void
tile_xy(int r, int ntx, int nty, int strip, int& bxi, int& byi)
{
    if (strip <= 0 || strip >= ntx) {
        bxi = r % ntx;
        byi = r / ntx;
        return;
    }
    const int s = r / (strip * nty);
    const int w = min(strip, ntx - s * strip);
    const int q = r - s * strip * nty;
    byi = q / w;
    bxi = s * strip + q % w;
}

// XCD-aware tile decode: consecutive hardware blocks go round-robin over the 8 XCDs; each XCD gets a contiguous range
// of tiles (in the block schedule's order, then z chunk) so neighbouring tiles share halo lines in the same L2.
__device__ __forceinline__ bool
decode_tile(int ntx, int nty, int ntz, int strip, int& bxi, int& byi, int& bzi)
{
    const int nt = ntx * nty * ntz;
    const int bid = blockIdx.x;
    const int per = (nt + 7) / 8;
    const int t = (bid % 8) * per + bid / 8;
    if (t >= nt)
        return false;
    bzi = t / (ntx * nty);
    tile_xy(t - bzi * ntx * nty, ntx, nty, strip, bxi, byi);
    return true;
}

// v moved across lanes by the DPP control CTRL, into the rows selected by ROWS (lanes of other rows keep v)
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float
dpp(float v)
{
    const int i = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(i, i, CTRL, ROWS, 0xf, false));
}

// Maximum over the 64 lanes of the wave, returned to every lane. DPP moves on the VALU (no LDS round trips): pairs
// and quads by quad_perm, 8 and 16 lanes by the half-row and row mirrors, then lane 15 of each row broadcast into
// rows 1 and 3 and lane 31 into rows 2 and 3, so lane 63 holds the maximum of all rows.
__device__ __forceinline__ float
wave_max(float v)
{
    v = fmaxf(v, dpp<0xb1>(v));       // quad_perm [1, 0, 3, 2]
    v = fmaxf(v, dpp<0x4e>(v));       // quad_perm [2, 3, 0, 1]
    v = fmaxf(v, dpp<0x141>(v));      // row_half_mirror
    v = fmaxf(v, dpp<0x140>(v));      // row_mirror
    v = fmaxf(v, dpp<0x142, 0xa>(v)); // row_bcast:15
    v = fmaxf(v, dpp<0x143, 0xc>(v)); // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Stores one 64-point row segment of a snapshot field (all 64 lanes of the wave call this). I = index type.
template <class I>
__device__ __forceinline__ void
snap_store(int mode, void* arr, float* scales, I idx, I sidx, bool act, float v)
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
        if (threadIdx.x == 0)
            scales[sidx] = ldexpf(1.0f, e - 14);
    } else if (mode == TTI_SNAP_BF16) {
        if (act)
            reinterpret_cast<__hip_bfloat16*>(arr)[idx] = __float2bfloat16(v);
    }
}

template <class I>
__device__ __forceinline__ float
snap_load(int mode, const void* arr, const float* scales, I idx, I sidx)
{
    if (mode == TTI_SNAP_FP32)
        return reinterpret_cast<const float*>(arr)[idx];
    if (mode == TTI_SNAP_FP16)
        return __half2float(reinterpret_cast<const __half*>(arr)[idx]) * scales[sidx];
    if (mode == TTI_SNAP_BF16)
        return __bfloat162float(reinterpret_cast<const __hip_bfloat16*>(arr)[idx]);
    return 0.0f;
}

// The stored words of one snapshot value, loaded one plane before they are used and decoded then by snap_decode
// (a decode right after the load would wait for that load, and for every memory operation issued before it)
struct SnapWord
{
    unsigned v; // FP32 bits, or the 16-bit FP16 / BF16 word
    float s;    // FP16: the segment's scale
};

template <class I>
__device__ __forceinline__ SnapWord
snap_fetch(int mode, const void* arr, const float* scales, I idx, I sidx)
{
    SnapWord r{ 0u, 1.0f };
    if (mode == TTI_SNAP_FP32) {
        r.v = reinterpret_cast<const unsigned*>(arr)[idx];
    } else if (mode == TTI_SNAP_FP16) {
        r.v = reinterpret_cast<const unsigned short*>(arr)[idx];
        r.s = scales[sidx];
    } else if (mode == TTI_SNAP_BF16) {
        r.v = reinterpret_cast<const unsigned short*>(arr)[idx];
    }
    return r;
}

// the value snap_load returns for the same words
__device__ __forceinline__ float
snap_decode(int mode, SnapWord r)
{
    if (mode == TTI_SNAP_FP32)
        return __uint_as_float(r.v);
    if (mode == TTI_SNAP_FP16)
        return __half2float(__ushort_as_half((unsigned short)r.v)) * r.s;
    if (mode == TTI_SNAP_BF16)
        return __uint_as_float(r.v << 16); // BF16 is the high half of the FP32 word
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
//   A  load the column words of plane o+1, store the prefetched tile of plane p+1 into buffer 1-cur, fetch the tile
//      of plane p+2 (loads are issued before this plane's stores: one counter tracks both)
//   B  staggered derivatives d_x, d_y of plane p on the tile rim -> LDS, barrier
//   C  D+ of d_x, d_y, the x-y mixed term and the z-mixed partial sums of plane p (registers)
//   D  d_z(o+4) from the column queue, finalise plane o from the pointwise values prefetched one plane earlier,
//      load the pointwise values of plane o+1, store plane o
// Two barriers per plane (between B and C, and at the end): tiles and derivative planes alternate between two LDS
// buffers by the parity of o.
// The loop body is instantiated for both parities of o (queue windows and derivative buffers are static).
// Global addresses are 32-bit byte offsets from the (scalar) base pointers: the launcher checks that every volume is
// smaller than 4 GiB.
// ------------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ float
ldf(const float* __restrict__ p, unsigned off)
{
    return *reinterpret_cast<const float*>(__builtin_assume_aligned(reinterpret_cast<const char*>(p) + off, 4));
}
__device__ __forceinline__ float4
ldf4(const float* __restrict__ p, unsigned off) // off = 0 mod 16
{
    return *reinterpret_cast<const float4*>(__builtin_assume_aligned(reinterpret_cast<const char*>(p) + off, 16));
}
__device__ __forceinline__ void
stf(float* __restrict__ p, unsigned off, float v)
{
    *reinterpret_cast<float*>(__builtin_assume_aligned(reinterpret_cast<char*>(p) + off, 4)) = v;
}

// Byte offset (within a plane) of the tile quad that starts at (gx, gy), clamped into the allocated rows
// [iybeg-4, iyend+4] and quad starts [ixbeg-4, ldx-4]. Clamped quads read real data of the volume that no output
// uses: values outside the allocated region only feed derivatives of non-interior points, which are zeroed.
__device__ __forceinline__ unsigned
quad_off(const Box& g, int gx, int gy)
{
    gx = min(max(gx, g.ixbeg - R), (int)g.ldx - 4);
    gy = min(max(gy, g.iybeg - R), g.iyend + R);
    return ((unsigned)gy * (unsigned)g.ldx + (unsigned)gx) * 4u;
}

// rho = 1/b by the hardware reciprocal (v_rcp_f32, 1 ulp) instead of an IEEE division sequence; zero where b is
// zero (outside the models' allocation)
__device__ __forceinline__ float
recip(float b)
{
    return b != 0.0f ? __builtin_amdgcn_rcpf(b) : 0.0f;
}

__device__ __forceinline__ bool
in_xy(const Box& g, int x, int y)
{
    return x >= g.ixbeg && x <= g.ixend && y >= g.iybeg && y <= g.iyend;
}

// staggered difference sum_n c_n (v[3+n] - v[4-n]) over 8 consecutive values v[0..7]
template <class F>
__device__ __forceinline__ float
stg8(F&& v)
{
    float a = 0.0f;
#pragma unroll
    for (int n = 1; n <= R; ++n)
        a += cstg(n) * (v(3 + n) - v(4 - n));
    return a;
}

#ifndef TTI_FWD_BY
#define TTI_FWD_BY 4
#endif
constexpr int FWD_BY = TTI_FWD_BY;         // rows per forward block (4: two 256-thread blocks per CU)
#ifndef TTI_FWD_UNROLL
#define TTI_FWD_UNROLL 2
#endif
#ifndef TTI_FWD_SETPRIO
#define TTI_FWD_SETPRIO 1
#endif
constexpr int FU = TTI_FWD_UNROLL;         // forward planes per loop iteration (2 or 6)
constexpr bool FWD_SETPRIO = TTI_FWD_SETPRIO != 0;
static_assert(FU == 2 || FU == 4 || FU == 6, "forward unroll of 2, 4 or 6 planes");
constexpr int XITEMS = BY * BW;            // d_x values per plane (576)
constexpr int XEXTRA = XITEMS - BX * BY;   // handled by a second item on the first threads (64)
static_assert(XEXTRA > 0 && XEXTRA <= BX * BY, "two d_x items per thread at most");
static_assert(BH * BX == 2 * BX * BY, "two d_y items per thread");

// One block's tile (bxi, byi, bzi) of a forward step; called by the single-volume kernel and the batched one.
// FBY = rows per block (8: one 512-thread block per CU; 4: two independent 256-thread blocks per CU).
template <int FBY, int SNAP, bool SPONGE, bool SIG = false>
__device__ __forceinline__ void
fwd_tile(const float* __restrict__ s1, const float* __restrict__ s2, float* __restrict__ s1o, float* __restrict__ s2o,
         float* __restrict__ u, float* __restrict__ w, const float* __restrict__ m1, const float* __restrict__ m2,
         const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ vrx,
         const float* __restrict__ vry, const float* __restrict__ vrz, const float* __restrict__ b,
         const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, const Box& g, float dt,
         float invdx, float invdy, float invdz, float i4xy, float i4yz, float i4zx, unsigned src_off, float src_val,
         const Snap& snap, int bz, int bxi, int byi, int bzi)
{
    // Tiles and derivative planes are double-buffered by the parity of o: a tile buffer is refilled by the next
    // plane's stage A, after the barrier that ends this plane.
    constexpr int NT = BX * FBY;                  // threads
    constexpr int FTH = FBY + 16, FBH = FBY + 8;  // tile heights (halo 8, halo 4)
    constexpr int FTQ = TQ_PER_ROW * FTH, FBQ = BQ_PER_ROW * FBH;
    constexpr int QT = (FTQ + NT - 1) / NT, QB = (FBQ + NT - 1) / NT; // quads per thread
    constexpr int FXI = FBY * BW, XI = (FXI + NT - 1) / NT;           // d_x items per thread
    constexpr int YI = FBH / FBY;                                     // d_y items per thread
    static_assert(FBH % FBY == 0, "d_y rows split evenly over the thread rows");
    __shared__ __align__(16) float T1[2][FTH][TW];
    __shared__ __align__(16) float T2[2][FTH][TW]; // sigma, halo 8: (x0-8+lx, y0-8+ly)
    __shared__ __align__(16) float Bt[2][FBH][BW]; // b, halo 4: (x0-4+lx, y0-4+ly)
    __shared__ float X1[2][FBY][BW], X2[2][FBY][BW]; // d_x at (x0-4+j, y0+r)
    __shared__ float Y1[2][FBH][BX], Y2[2][FBH][BX]; // d_y at (x0+c, y0-4+j)

    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * BX + tx;
    const int x0 = g.ixbeg + bxi * BX, y0 = g.iybeg + byi * FBY, z0 = g.izbeg + bzi * bz;
    const int z1 = min(z0 + bz, g.izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= g.ixend && iy <= g.iyend;
    const unsigned PB = (unsigned)g.ldx * (unsigned)g.ldy * 4u; // bytes per plane
    const unsigned colb = active ? ((unsigned)iy * (unsigned)g.ldx + (unsigned)ix) * 4u : 0u;
    // snapshot element offsets (32-bit: the launchers check that a snapshot field has < 2^32 elements): plane
    // stride, this thread's column within a plane, and its 64-point segment's scale within a plane
    const unsigned s_plane = (unsigned)(snap.ldy * snap.ldx);
    const unsigned s_col = (unsigned)(iy - g.iybeg) * (unsigned)snap.ldx + (unsigned)(ix - g.ixbeg);
    const unsigned s_scol = (unsigned)(iy - g.iybeg) * (unsigned)(snap.ldx / SNAP_SEG) + (unsigned)((x0 - g.ixbeg) / SNAP_SEG);
    float axy = 1.0f, bxy = 1.0f;
    if (SPONGE && active) {
        axy = fminf(abx[ix].a, aby[iy].a);
        bxy = fminf(abx[ix].b, aby[iy].b);
    }
    // tile quads (clamped into the allocation, see quad_off); slots beyond the count map to the last quad (their
    // loads and stores are skipped)
    int tq_row[QT], tq_lx[QT], bq_row[QB], bq_lx[QB];
    unsigned tqb[QT], bqb[QB];
#pragma unroll
    for (int k = 0; k < QT; ++k) {
        const int q = min(tid + k * NT, FTQ - 1);
        tq_row[k] = q / TQ_PER_ROW;
        tq_lx[k] = 4 * (q - tq_row[k] * TQ_PER_ROW);
        tqb[k] = quad_off(g, x0 - 8 + tq_lx[k], y0 - 8 + tq_row[k]);
    }
#pragma unroll
    for (int k = 0; k < QB; ++k) {
        const int q = min(tid + k * NT, FBQ - 1);
        bq_row[k] = q / BQ_PER_ROW;
        bq_lx[k] = 4 * (q - bq_row[k] * BQ_PER_ROW);
        bqb[k] = quad_off(g, x0 - 4 + bq_lx[k], y0 - 4 + bq_row[k]);
    }
    // Prefetch registers come in two sets selected by the parity of the plane: a body fills the set the next body
    // drains, so no loaded register is ever copied (a copy would wait for the load, and for every load before it).
    float4 pt1[2][QT], pt2[2][QT], pb[2][QB];
    // quad slots past a tile's quad count load and store nothing, so whole waves skip them (at 64 x 4 the second
    // sigma slot holds 144 quads for 256 threads: 2.98 instead of 3.04 ms per step at 512^3)
    auto tq_ok = [&](int k) { return k * NT + NT <= FTQ || tid + k * NT < FTQ; };
    auto bq_ok = [&](int k) { return k * NT + NT <= FBQ || tid + k * NT < FBQ; };
    auto fetch = [&](int pp, int set) {
        const unsigned o = (unsigned)pp * PB;
#pragma unroll
        for (int k = 0; k < QT; ++k) {
            if (tq_ok(k)) {
                pt1[set][k] = ldf4(s1, o + tqb[k]);
                pt2[set][k] = ldf4(s2, o + tqb[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < QB; ++k)
            if (bq_ok(k))
                pb[set][k] = ldf4(b, o + bqb[k]);
    };
    auto commit = [&](int buf, int set) {
#pragma unroll
        for (int k = 0; k < QT; ++k) {
            if (tq_ok(k)) {
                *reinterpret_cast<float4*>(&T1[buf][tq_row[k]][tq_lx[k]]) = pt1[set][k];
                *reinterpret_cast<float4*>(&T2[buf][tq_row[k]][tq_lx[k]]) = pt2[set][k];
            }
        }
#pragma unroll
        for (int k = 0; k < QB; ++k)
            if (bq_ok(k))
                *reinterpret_cast<float4*>(&Bt[buf][bq_row[k]][bq_lx[k]]) = pb[set][k];
    };
    // d_x items (r, j): item k is tid + k * NT; d_y items: rows ty + h * FBY of the halo-4 band, column tx.
    // The scale is 1/d for interior points and zero elsewhere (floats: no lane masks live across the loop).
    int xr[XI], xj[XI];
    float xs[XI], ys[YI];
#pragma unroll
    for (int k = 0; k < XI; ++k) {
        const int it = min(tid + k * NT, FXI - 1);
        xr[k] = it / BW;
        xj[k] = it - xr[k] * BW;
        xs[k] = in_xy(g, x0 - 4 + xj[k], y0 + xr[k]) ? invdx : 0.0f; // repeated items compute the same value
    }
#pragma unroll
    for (int h = 0; h < YI; ++h)
        ys[h] = in_xy(g, x0 + tx, y0 - 4 + ty + h * FBY) ? invdy : 0.0f;

    // Register queues along z. The loop body is instantiated for FU consecutive planes; the body of plane o with
    // c = o - (z0 - 8) mod FU sees each queue through a window starting at slot c (FU - 1 spare slots per queue), and
    // after the FU-th plane the queues shift by FU: one register move per live value every FU planes.
    float sq1[8 + FU], sq2[8 + FU]; // sigma(o + k)   at [c + k], k = 0..7 (slot c + 8 receives sigma(o + 8))
    float dz1[7 + FU], dz2[7 + FU]; // d_z(o - 3 + k) at [c + k], k = 0..7
    float xx1[4 + FU], xx2[4 + FU]; // D+_x d_x (o + k), unscaled by rho, at [c + k], k = 0..4
    float yy1[4 + FU], yy2[4 + FU]; // D+_y d_y
    float xy1[4 + FU], xy2[4 + FU]; // X_xy sigma
    float qx1[8 + FU], qx2[8 + FU]; // z-x mixed partial sums of output o + j at [c + j], j = 0..8
    float qy1[8 + FU], qy2[8 + FU]; // y-z mixed partial sums
#pragma unroll
    for (int k = 0; k < 7 + FU; ++k)
        dz1[k] = dz2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4 + FU; ++k)
        xx1[k] = xx2[k] = yy1[k] = yy2[k] = xy1[k] = xy2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8 + FU; ++k)
        qx1[k] = qx2[k] = qy1[k] = qy2[k] = 0.0f;
    // column words: sigma and b at a plane clamped into the allocation (a clamped sigma only meets a zero b; b is
    // zeroed outside the interior planes where it is used, not where it is loaded)
    auto col = [&](const float* __restrict__ v, int q) {
        return ldf(v, (unsigned)min(max(q, g.izbeg - R), g.izend + R) * PB + colb);
    };
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        sq1[k] = col(s1, z0 - 8 + k);
        sq2[k] = col(s2, z0 - 8 + k);
    }
    // b(o + k) at [c + k], k = 0..5: b(o + 4) for d_z(o + 4), and b(o) for the output plane, which is not loaded
    // again 5 planes later when L2 no longer holds it (2.95 -> 2.87 ms per step at 512^3, 3.11 -> 3.03 ms with the
    // FP16 snapshot)
    float bq[5 + FU];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        bq[k] = col(b, z0 - 8 + k);
    // pointwise words of an output plane, one set per parity: m1 m2 m3 m4 rx ry rz, slot 7 unused (b comes from the
    // b queue), u w
    float pw[2][10];
    auto fetch_point = [&](int o, int set) {
        if (o >= z0 && o < z1) {
            const unsigned io = (unsigned)o * PB + colb;
            pw[set][0] = ldf(m1, io);
            pw[set][1] = ldf(m2, io);
            pw[set][2] = ldf(m3, io);
            pw[set][3] = ldf(m4, io);
            pw[set][4] = ldf(vrx, io);
            pw[set][5] = ldf(vry, io);
            pw[set][6] = ldf(vrz, io);
            pw[set][8] = ldf(u, io);
            pw[set][9] = ldf(w, io);
        }
    };

    fetch(z0 - R, 0);
    commit(0, 0);
    fetch(z0 - R + 1, 0);
    __syncthreads();

    auto body = [&](auto curc, int o) {
        constexpr int cur = decltype(curc)::value; // o - (z0 - 8) mod FU: queue window
        constexpr int par = cur & 1;               // parity of o: prefetch set and derivative buffer
        // tile buffer of plane p and of plane p + 1, by parity: compile-time, so each LDS address is a thread base
        // plus constants (with three buffers rotating at run time under a two-plane unroll, every tap address took
        // a scalar buffer base: 45 more VALU instructions per two planes; the second barrier per plane that two
        // buffers need costs less: 2.90 -> 2.88 ms per step at 512^3, 3.06 -> 3.02 ms with the FP16 snapshot)
        constexpr int tb = par, tn = par ^ 1;
        const int p = o + R;
        const float pin = (p >= g.izbeg && p <= g.izend) ? 1.0f : 0.0f;
        const int qd = o + R;
        const float bd = (qd >= g.izbeg && qd <= g.izend) ? bq[cur + 4] : 0.0f;
        // ---- A: issue this plane's loads first, so that waiting for them never waits for the stores below (one
        // counter covers loads and stores): column words of plane o + 1, then the tile of plane p + 1 goes to buffer
        // tn (last read by the previous plane, before the barrier that ended it) and the tile of plane p + 2 is fetched. The
        // wave runs this stage at raised priority, so its loads go out ahead of the other block's arithmetic
        // (measured at 512^3: -1% per step, -2% with the FP16 snapshot; the adjoint gains nothing from it).
        if (FWD_SETPRIO)
            __builtin_amdgcn_s_setprio(3);
        sq1[cur + 8] = col(s1, o + 8);
        sq2[cur + 8] = col(s2, o + 8);
        bq[cur + 5] = col(b, o + 5);
        if (p + 1 <= z1 + R - 1)
            commit(tn, par);
        if (p + 2 <= z1 + R - 1)
            fetch(p + 2, par ^ 1);
        // pointwise words of plane o + 1 into the set that plane o - 1 drained in its stage D: issued here they have
        // the whole plane to arrive (at the end of stage D they had part of the next one: 3.34 -> 3.14 ms for the
        // FP16-snapshot step at 512^3, 3.06 -> 3.04 ms without; the adjoint's pointwise words gain nothing from it)
        fetch_point(o + 1, par ^ 1);
        if (FWD_SETPRIO)
            __builtin_amdgcn_s_setprio(0);
        // ---- B: d_x on rows of the tile (x' in [x0-4, x0+68)), d_y on columns (y' in [y0-4, y0+BY+4))
        // (the scale is zero for non-interior points: every tile value is finite)
#pragma unroll
        for (int i = 0; i < XI; ++i) {
            // items past the count (first wave only for the last slot) repeat the last item: same value, same word
            if (i == 0 || tid + i * NT < FXI || tid < BX) {
                const float bb = Bt[tb][xr[i] + 4][xj[i]] * (xs[i] * pin);
                X1[par][xr[i]][xj[i]] = bb * stg8([&](int k) { return T1[tb][xr[i] + 8][xj[i] + k]; });
                X2[par][xr[i]][xj[i]] = bb * stg8([&](int k) { return T2[tb][xr[i] + 8][xj[i] + k]; });
            }
        }
#pragma unroll
        for (int h = 0; h < YI; ++h) {
            const int j = ty + h * FBY;
            const float bb = Bt[tb][j][tx + 4] * (ys[h] * pin);
            Y1[par][j][tx] = bb * stg8([&](int k) { return T1[tb][j + k][tx + 8]; });
            Y2[par][j][tx] = bb * stg8([&](int k) { return T2[tb][j + k][tx + 8]; });
        }
        __syncthreads();
        // ---- C: plane-p contributions
        if (active) {
            float fx1 = 0, fx2 = 0, fy1 = 0, fy2 = 0, fxy1 = 0, fxy2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                fx1 += c * (X1[par][ty][tx + 4 + n] - X1[par][ty][tx + 5 - n]);
                fx2 += c * (X2[par][ty][tx + 4 + n] - X2[par][ty][tx + 5 - n]);
                fy1 += c * (Y1[par][ty + 4 + n][tx] - Y1[par][ty + 5 - n][tx]);
                fy2 += c * (Y2[par][ty + 4 + n][tx] - Y2[par][ty + 5 - n][tx]);
                const float e = ecen(n);
                fxy1 += e * ((T1[tb][ty + 8 + n][tx + 8 + n] - T1[tb][ty + 8 - n][tx + 8 + n]) +
                             (T1[tb][ty + 8 - n][tx + 8 - n] - T1[tb][ty + 8 + n][tx + 8 - n]));
                fxy2 += e * ((T2[tb][ty + 8 + n][tx + 8 + n] - T2[tb][ty + 8 - n][tx + 8 + n]) +
                             (T2[tb][ty + 8 - n][tx + 8 - n] - T2[tb][ty + 8 + n][tx + 8 - n]));
                // plane p at x-offset +n feeds output p - n (q[4 - n]); at -n output p + n (q[4 + n]), sign flipped
                const float dx1 = T1[tb][ty + 8][tx + 8 + n] - T1[tb][ty + 8][tx + 8 - n];
                const float dx2 = T2[tb][ty + 8][tx + 8 + n] - T2[tb][ty + 8][tx + 8 - n];
                const float dy1 = T1[tb][ty + 8 + n][tx + 8] - T1[tb][ty + 8 - n][tx + 8];
                const float dy2 = T2[tb][ty + 8 + n][tx + 8] - T2[tb][ty + 8 - n][tx + 8];
                qx1[cur + 4 - n] += e * dx1;
                qx1[cur + 4 + n] -= e * dx1;
                qx2[cur + 4 - n] += e * dx2;
                qx2[cur + 4 + n] -= e * dx2;
                qy1[cur + 4 - n] += e * dy1;
                qy1[cur + 4 + n] -= e * dy1;
                qy2[cur + 4 - n] += e * dy2;
                qy2[cur + 4 + n] -= e * dy2;
            }
            xx1[cur + 4] = fx1 * invdx;
            xx2[cur + 4] = fx2 * invdx;
            yy1[cur + 4] = fy1 * invdy;
            yy2[cur + 4] = fy2 * invdy;
            xy1[cur + 4] = fxy1 * i4xy;
            xy2[cur + 4] = fxy2 * i4xy;
        }
        // ---- D: d_z(o+4) and the finalisation of plane o
        {
            const float dzs = bd * invdz;
            dz1[cur + 7] = dzs * stg8([&](int k) { return sq1[cur + k]; });
            dz2[cur + 7] = dzs * stg8([&](int k) { return sq2[cur + k]; });
            if (g.fs && o == g.izbeg) {
                mirror_dz<cur>(dz1);
                mirror_dz<cur>(dz2);
            }
        }
        const bool fin = o >= z0 && active;
        float unew = 0.0f, wnew = 0.0f, s1n = 0.0f, s2n = 0.0f;
        const unsigned io = (unsigned)o * PB + colb;
        if (fin) {
            const float fz1 = stg8([&](int k) { return dz1[cur + k]; });
            const float fz2 = stg8([&](int k) { return dz2[cur + k]; });
            const float* P = pw[par];
            const float mm1 = P[0], mm2 = P[1], mm3 = P[2], mm4 = P[3], vx = P[4], vy = P[5], vz = P[6];
            const float uo = P[8], wo = P[9];
            const float rho = recip(bq[cur]);
            const float f1xx = xx1[cur] * rho, f2xx = xx2[cur] * rho;
            const float f1yy = yy1[cur] * rho, f2yy = yy2[cur] * rho;
            const float f1zz = fz1 * invdz * rho, f2zz = fz2 * invdz * rho;
            const float f1xy = xy1[cur], f2xy = xy2[cur];
            const float f1zx = qx1[cur] * i4zx, f2zx = qx2[cur] * i4zx;
            const float f1yz = qy1[cur] * i4yz, f2yz = qy2[cur] * i4yz;
            const float rx2 = vx * vx, ry2 = vy * vy, rz2 = vz * vz;
            const float rxy = 2.0f * vx * vy, ryz = 2.0f * vy * vz, rzx = 2.0f * vz * vx;
            // sums of products are written as explicit fmaf chains (left to right): with contraction left to the
            // compiler, copies of this body (the loop's warm-up planes vs the rest) may fuse different products and
            // round differently, which would make a point's value depend on where its z chunk starts
            const float V1 = fmaf(ryz, f1yz, fmaf(rzx, f1zx, fmaf(rxy, f1xy, fmaf(rz2, f1zz, fmaf(ry2, f1yy, rx2 * f1xx)))));
            const float H1 = __fsub_rn(__fadd_rn(__fadd_rn(f1xx, f1yy), f1zz), V1); // plain adds: no product fused
            const float V2 = fmaf(ryz, f2yz, fmaf(rzx, f2zx, fmaf(rxy, f2xy, fmaf(rz2, f2zz, fmaf(ry2, f2yy, rx2 * f2xx)))));
            const float H2 = __fsub_rn(__fadd_rn(__fadd_rn(f2xx, f2yy), f2zz), V2);
            const float r3 = fmaf(mm2, V2, fmaf(mm4, V1, mm1 * H1));
            const float r4 = fmaf(mm4, H2, fmaf(mm3, V2, mm2 * H1));
            unew = uo + r3 * dt;
            wnew = wo + r4 * dt;
            if (io == src_off) {
                unew += src_val;
                wnew += src_val;
            }
            float a = 1.0f, bb = 1.0f;
            if constexpr (SPONGE) {
                a = fminf(axy, abz[o].a);
                bb = fminf(bxy, abz[o].b);
            }
            s1n = fmaf(bb * unew, dt, a * sq1[cur]);
            s2n = fmaf(bb * wnew, dt, a * sq2[cur]);
        }
        // the stores of plane o
        if (fin) {
            stf(u, io, unew);
            stf(w, io, wnew);
            stf(s1o, io, s1n);
            stf(s2o, io, s2n);
        }
        if (SNAP != TTI_SNAP_NONE && o >= z0 && iy <= g.iyend) {
            const unsigned pz = (unsigned)(o - g.izbeg);
            const unsigned idx = pz * s_plane + s_col;
            const unsigned sidx = pz * (s_plane / SNAP_SEG) + s_scol;
            snap_store(SNAP, snap.a, snap.sa, idx, sidx, active, unew);
            snap_store(SNAP, snap.b, snap.sb, idx, sidx, active, wnew);
            if constexpr (SIG) { // the sigma this step read (its u, w update used it)
                snap_store(SNAP, snap.c, snap.sc, idx, sidx, active, sq1[cur]);
                snap_store(SNAP, snap.d, snap.sd, idx, sidx, active, sq2[cur]);
            }
        }
        __syncthreads(); // tile buffer tb is refilled by the next plane
        if (cur == FU - 1) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                sq1[k] = sq1[k + FU];
                sq2[k] = sq2[k + FU];
            }
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                dz1[k] = dz1[k + FU];
                dz2[k] = dz2[k + FU];
            }
#pragma unroll
            for (int k = 0; k < 5; ++k)
                bq[k] = bq[k + FU];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                xx1[k] = xx1[k + FU];
                xx2[k] = xx2[k + FU];
                yy1[k] = yy1[k + FU];
                yy2[k] = yy2[k + FU];
                xy1[k] = xy1[k + FU];
                xy2[k] = xy2[k + FU];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                qx1[k] = qx1[k + FU];
                qx2[k] = qx2[k + FU];
                qy1[k] = qy1[k + FU];
                qy2[k] = qy2[k + FU];
            }
#pragma unroll
            for (int k = 8; k < 8 + FU; ++k)
                qx1[k] = qx2[k] = qy1[k] = qy2[k] = 0.0f;
        }
    };

    // FU bodies per iteration; the chunk may end after any of them
    for (int o = z0 - 8; o < z1; o += FU) {
        body(std::integral_constant<int, 0>{}, o);
        if (o + 1 >= z1)
            break;
        body(std::integral_constant<int, 1>{}, o + 1);
        if constexpr (FU >= 4) {
            if (o + 2 >= z1)
                break;
            body(std::integral_constant<int, 2>{}, o + 2);
            if (o + 3 >= z1)
                break;
            body(std::integral_constant<int, 3>{}, o + 3);
        }
        if constexpr (FU == 6) {
            if (o + 4 >= z1)
                break;
            body(std::integral_constant<int, 4>{}, o + 4);
            if (o + 5 >= z1)
                break;
            body(std::integral_constant<int, 5>{}, o + 5);
        }
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
    // explicit fmaf: one fixed rounding for every copy of the code (see the forward finalisation)
    t.lap1 = fmaf(mm2, f2, mm1 * f1);
    t.hz1 = fmaf(-mm2, f2, (-mm1 + mm4) * f1);
    t.lap2 = mm4 * f2;
    t.hz2 = fmaf(-mm4 + mm3, f2, mm2 * f1);
    return t;
}

// Adjoint tile shape: AX x AY output columns per block, one per thread (AX = 64: a wave is one 64-point row;
// AX = 32: a wave is two 32-point rows, each half-wave one contiguous LDS row). AY rows.
#ifndef TTI_ADJ_AX
#define TTI_ADJ_AX 64
#endif
#ifndef TTI_ADJ_AY
#define TTI_ADJ_AY 8
#endif
constexpr int ADJ_AX = TTI_ADJ_AX, ADJ_AY = TTI_ADJ_AY;

template <int AX, int AY>
struct AdjShape
{
    static constexpr int NT = AX * AY;                 // threads
    static constexpr int TW = AX + 16, TH = AY + 16;   // footprint with halo 8: (x0-8+lx, y0-8+ly)
    static constexpr int BW = AX + 8, BH = AY + 8;     // halo-4 region
    static constexpr int QR = TW / 4;                  // quads per footprint row
    // quads of the plus-shaped footprint: rows [8, 8+AY) whole; rows [4, 8) and [8+AY, 12+AY) from x' = 4 (the halo-4
    // square); the 4 outermost rows on each side from x' = 8 (the y halo of the block's own columns)
    static constexpr int NQ = AY * QR + 8 * (QR - 2) + 8 * (QR - 4);
    static constexpr int FXI = AY * BW, XI = (FXI + NT - 1) / NT; // d_x items and items per thread
    static constexpr int YI = (BH + AY - 1) / AY;                  // d_y items per thread
    static_assert(AX % 32 == 0 && AX <= 64 && NT % 64 == 0 && SNAP_SEG % AX == 0, "adjoint tile shape");
    static_assert(NQ <= NT, "one footprint quad per thread");
};

template <int AX, int AY>
struct __align__(16) AdjTiles
{
    using Sh = AdjShape<AX, AY>;
    // the two fields of each temporary interleaved: every stencil tap is one 8-byte LDS read for both (measured at
    // 512^3: 3.87 -> 3.75 ms per step, 4.42 -> 4.31 ms with imaging; the forward's sigma tiles lost 2.5% the same way)
    float2 SX[AY][Sh::TW];      // S_x at (x0-8+lx, y0+r)
    float2 SY[Sh::TH][AX];      // S_y at (x0+c, y0-8+ly)
    float2 MXY[Sh::BH][Sh::BW]; // M_xy at (x0-4+lx, y0-4+ly)
    float2 MZX[AY][Sh::BW];     // M_zx at (x0-4+lx, y0+r)
    float2 MYZ[Sh::BH][AX];     // M_yz at (x0+c, y0-4+ly)
    float Bt[Sh::BH][Sh::BW];                   // b at (x0-4+lx, y0-4+ly)
};

// Footprint quad t (t < NQ) of the plus-shaped adjoint footprint: its row and x' offset, bands of rows in order
template <int AX, int AY>
__device__ __forceinline__ void
adj_quad(int t, int& row, int& lx)
{
    using Sh = AdjShape<AX, AY>;
    constexpr int QR = Sh::QR;
    // (first row, rows, quads per row, first quad) of the five bands, the two outer ones (S_y only) first, so that
    // the threads of those bands fill whole waves
    constexpr int rs[5] = { 0, 12 + AY, 4, 8, 8 + AY }, br[5] = { 4, 4, 4, AY, 4 };
    constexpr int bq[5] = { QR - 4, QR - 4, QR - 2, QR, QR - 2 }, b0[5] = { 2, 2, 1, 0, 1 };
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const int n = br[k] * bq[k];
        if (t < n) {
            row = rs[k] + t / bq[k];
            lx = 4 * (b0[k] + t % bq[k]);
            return;
        }
        t -= n;
    }
    row = 0;
    lx = 0;
}

// One block's tile of an adjoint step; SNAP = snapshot mode of the imaging condition (TTI_SNAP_NONE: no imaging)
template <int SNAP, bool SPONGE, int AX = ADJ_AX, int AY = ADJ_AY>
__device__ __forceinline__ void
adj_tile(const float* __restrict__ a1, const float* __restrict__ a2, float* __restrict__ a1o, float* __restrict__ a2o,
         float* __restrict__ u, float* __restrict__ w, const float* __restrict__ m1, const float* __restrict__ m2,
         const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ vrx,
         const float* __restrict__ vry, const float* __restrict__ vrz, const float* __restrict__ b,
         const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, const Box& g, float dt,
         float invdx, float invdy, float invdz, float i4xy, float i4yz, float i4zx, const Snap& snap,
         double* __restrict__ grad, int bz, int bxi, int byi, int bzi)
{
    using Sh = AdjShape<AX, AY>;
    constexpr int NT = Sh::NT, BW = Sh::BW, BH = Sh::BH, XI = Sh::XI, YI = Sh::YI, FXI = Sh::FXI;
    // planes per loop iteration: the z queues shift once per iteration, and the shift is a register move per live
    // value (70 of the 520 VALU instructions per plane at 2 planes, 49 of 498 at 4). Four planes where the registers
    // allow it (512^3: 3.575 -> 3.558 ms per step, same bits); with imaging they spill 4 VGPRs (4.06 -> 4.27 ms)
    constexpr int AU = SNAP == TTI_SNAP_NONE ? 4 : 2;
    __shared__ AdjTiles<AX, AY> S[2];
    __shared__ float2 X[AY][BW]; // d_x of both fields
    __shared__ float2 Y[BH][AX]; // d_y

    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * AX + tx;
    const int x0 = g.ixbeg + bxi * AX, y0 = g.iybeg + byi * AY, z0 = g.izbeg + bzi * bz;
    const int z1 = min(z0 + bz, g.izend + 1);
    const int ix = x0 + tx, iy = y0 + ty;
    const bool active = ix <= g.ixend && iy <= g.iyend;
    const unsigned PB = (unsigned)g.ldx * (unsigned)g.ldy * 4u;
    const unsigned colb = active ? ((unsigned)iy * (unsigned)g.ldx + (unsigned)ix) * 4u : 0u;
    float axy = 1.0f, bxy = 1.0f;
    if (SPONGE && active) {
        axy = fminf(abx[ix].a, aby[iy].a);
        bxy = fminf(abx[ix].b, aby[iy].b);
    }
    // snapshot and image element offsets (32-bit: the launchers check both fit): plane strides and this thread's
    // column within a plane; the block lies inside one 64-point segment of a snapshot row
    const unsigned s_plane = (unsigned)(snap.ldy * snap.ldx);
    const unsigned s_col = (unsigned)(iy - g.iybeg) * (unsigned)snap.ldx + (unsigned)(ix - g.ixbeg);
    const unsigned s_scol =
        (unsigned)(iy - g.iybeg) * (unsigned)(snap.ldx / SNAP_SEG) + (unsigned)((x0 - g.ixbeg) / SNAP_SEG);
    const unsigned g_row = (unsigned)(g.ixend - g.ixbeg + 1);
    const unsigned g_plane = g_row * (unsigned)(g.iyend - g.iybeg + 1);
    const unsigned g_col = (unsigned)(iy - g.iybeg) * g_row + (unsigned)(ix - g.ixbeg);

    // this thread's quad of the plus-shaped footprint (quads are 4-aligned in lx, as are the footprint regions);
    // threads beyond the quad count load nothing
    int tq_row = 0, tq_lx = 0;
    const bool tq_need = tid < Sh::NQ;
    // The 4 outer y rows on each side of the footprint (the first two bands, whole waves at AX = 64 and 32) need S_y
    // only: their threads load 8 words (no rx, rz) and form nothing else (3.76 -> 3.66 ms per step at 512^3)
    const bool sy_only = tid < 8 * (Sh::QR - 4);
    adj_quad<AX, AY>(tq_need ? tid : 0, tq_row, tq_lx);
    const bool row8 = tq_row >= 8 && tq_row < 8 + AY;
    const bool col8 = tq_lx >= 8 && tq_lx < 8 + AX;
    const bool sq4 = tq_lx >= 4 && tq_lx < 4 + BW && tq_row >= 4 && tq_row < 4 + BH;
    const unsigned tqb = quad_off(g, x0 - 8 + tq_lx, y0 - 8 + tq_row); // clamped, see quad_off
    float4 qa1, qa2, qm1, qm2, qm3, qm4, qrx, qry, qrz, qb;
    auto fetch = [&](int pp) {
        if (!tq_need)
            return;
        const unsigned i = (unsigned)pp * PB + tqb;
        qa1 = ldf4(a1, i);
        qa2 = ldf4(a2, i);
        qm1 = ldf4(m1, i);
        qm2 = ldf4(m2, i);
        qm3 = ldf4(m3, i);
        qm4 = ldf4(m4, i);
        qry = ldf4(vry, i);
        qb = ldf4(b, i);
        if (!sy_only) {
            qrx = ldf4(vrx, i);
            qrz = ldf4(vrz, i);
        }
    };
    auto commit = [&](int buf) {
        if (!tq_need)
            return;
        AdjTiles<AX, AY>& t = S[buf];
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
        auto st2 = [](float2* dst, const float* a, const float* b) { // 4 points of two fields, interleaved
            *reinterpret_cast<float4*>(dst) = make_float4(a[0], b[0], a[1], b[1]);
            *reinterpret_cast<float4*>(dst + 2) = make_float4(a[2], b[2], a[3], b[3]);
        };
        if (sy_only) {
            float y1[4], y2[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const Temps tt = temps(A1[e], A2[e], M1[e], M2[e], M3[e], M4[e]);
                const float rho = recip(BB[e]);
                const float vy = RY[e];
                y1[e] = rho * (tt.lap1 + vy * vy * tt.hz1);
                y2[e] = rho * (tt.lap2 + vy * vy * tt.hz2);
            }
            st2(&t.SY[tq_row][tq_lx - 8], y1, y2);
            return;
        }
        float sx1[4], sx2[4], sy1[4], sy2[4], mxy1[4], mxy2[4], mzx1[4], mzx2[4], myz1[4], myz2[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const Temps tt = temps(A1[e], A2[e], M1[e], M2[e], M3[e], M4[e]);
            const float rho = recip(BB[e]);
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
            st2(&t.SX[tq_row - 8][tq_lx], sx1, sx2);
        }
        if (col8) {
            st2(&t.SY[tq_row][tq_lx - 8], sy1, sy2);
        }
        if (sq4) {
            st2(&t.MXY[tq_row - 4][tq_lx - 4], mxy1, mxy2);
            st(&t.Bt[tq_row - 4][tq_lx - 4], BB);
            if (row8) {
                st2(&t.MZX[tq_row - 8][tq_lx - 4], mzx1, mzx2);
            }
            if (col8) {
                st2(&t.MYZ[tq_row - 4][tq_lx - 8], myz1, myz2);
            }
        }
    };
    // d_x items (r, j) = item tid + i NT of the AY x BW rim rows; d_y items: rows ty + h AY of the BH band, column tx.
    // The scale is 1/d for interior points and zero elsewhere (floats: no lane masks live across the loop).
    int xr[XI], xj[XI];
    float xs[XI], ys[YI];
#pragma unroll
    for (int i = 0; i < XI; ++i) {
        const int it = min(tid + i * NT, FXI - 1); // items past the count repeat the last one (same value, same word)
        xr[i] = it / BW;
        xj[i] = it - xr[i] * BW;
        xs[i] = in_xy(g, x0 - 4 + xj[i], y0 + xr[i]) ? invdx : 0.0f;
    }
#pragma unroll
    for (int h = 0; h < YI; ++h)
        ys[h] = ty + h * AY < BH && in_xy(g, x0 + tx, y0 - 4 + ty + h * AY) ? invdy : 0.0f;

    // z queues seen through a window at slot c = o - z0 mod AU, shifted by AU every AU-th plane (as in the
    // forward kernel)
    float sz1[7 + AU], sz2[7 + AU];       // S_z(o + k)                       at [c + k], k = 0..7
    float dz1[7 + AU], dz2[7 + AU];       // d_z(o - 3 + k)                   at [c + k], k = 0..7
    // right-hand side partial sums of output o + j at [c + j], j = 0..8: the X_zx + X_yz terms of the planes
    // o + j - 4 .. o + j + 4 and, when plane o + j itself is on the tile, its in-plane terms D+_x d_x + D+_y d_y +
    // X_xy M_xy (no separate queue: they are summed into the same slot)
    float q1[8 + AU], q2[8 + AU];
#pragma unroll
    for (int k = 0; k < 7 + AU; ++k)
        dz1[k] = dz2[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 8 + AU; ++k)
        q1[k] = q2[k] = 0.0f;

    // S_z of the column at plane q, clamped into the allocation (a clamped value only meets a zero b below); the 8
    // column words are loaded one plane before they are combined, so the loads overlap a whole plane
    struct Col
    {
        float a1, a2, m1, m2, m3, m4, b, vz;
    };
    auto col_load = [&](int q) {
        const unsigned i = (unsigned)min(max(q, g.izbeg - R), g.izend + R) * PB + colb;
        return Col{ ldf(a1, i), ldf(a2, i), ldf(m1, i), ldf(m2, i), ldf(m3, i), ldf(m4, i), ldf(b, i), ldf(vrz, i) };
    };
    auto col_sz = [&](const Col& c, float& v1, float& v2) {
        const Temps t = temps(c.a1, c.a2, c.m1, c.m2, c.m3, c.m4);
        const float rho = recip(c.b);
        v1 = rho * (t.lap1 + c.vz * c.vz * t.hz1);
        v2 = rho * (t.lap2 + c.vz * c.vz * t.hz2);
    };
    // the column's sigma and b words ride along in queues of their own instead of being loaded again planes later
    // (when L2 no longer holds them): sigma of plane o + k at [c + k] (k = 0..7, the output plane's sigma for its
    // time update), b of plane o + 4 + k at [c + k] (k = 0..3, for d_z(o + 4)); 3.64 -> 3.57 ms per step at 512^3,
    // 4.18 -> 4.06 ms with imaging. b is zeroed outside the interior planes where it is used, not where it is
    // loaded: a select right after a load waits for that load and every load issued before it
    float sa1[7 + AU], sa2[7 + AU], sb[3 + AU];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const Col c = col_load(z0 - 8 + k);
        col_sz(c, sz1[k], sz2[k]);
        sa1[k] = c.a1;
        sa2[k] = c.a2;
        if (k >= 4)
            sb[k - 4] = c.b;
    }
    // prefetch registers in two sets selected by the parity of the plane (a body fills the set the next drains)
    Col ncol[2];
    ncol[0] = col_load(z0 - 1);
    // pointwise words of an output plane: u w (+ snapshot u, w and the image value)
    float pw[2][2];
    SnapWord sw_u[2], sw_w[2];
    double g0[2] = { 0.0, 0.0 };
    unsigned gi[2] = { 0, 0 };
    constexpr bool imaging = SNAP != TTI_SNAP_NONE;
    auto fetch_point = [&](int o, int set) {
        if (o >= z0 && o < z1) {
            const unsigned io = (unsigned)o * PB + colb;
            pw[set][0] = ldf(u, io);
            pw[set][1] = ldf(w, io);
            if (imaging && active) {
                const unsigned pz = (unsigned)(o - g.izbeg);
                const unsigned idx = pz * s_plane + s_col;
                const unsigned sidx = pz * (s_plane / SNAP_SEG) + s_scol;
                sw_u[set] = snap_fetch(SNAP, snap.a, snap.sa, idx, sidx);
                sw_w[set] = snap_fetch(SNAP, snap.b, snap.sb, idx, sidx);
                gi[set] = pz * g_plane + g_col;
                g0[set] = grad[gi[set]];
            }
        }
    };

    fetch(z0 - R);
    commit(0);
    fetch(z0 - R + 1);
    __syncthreads();

    auto body = [&](auto curc, int o) {
        constexpr int cur = decltype(curc)::value; // queue window
        constexpr int par = cur & 1;               // prefetch set and tile buffer
        const int p = o + R;
        const float pin = (p >= g.izbeg && p <= g.izend) ? 1.0f : 0.0f;
        col_sz(ncol[par], sz1[cur + 7], sz2[cur + 7]);
        sa1[cur + 7] = ncol[par].a1;
        sa2[cur + 7] = ncol[par].a2;
        sb[cur + 3] = ncol[par].b;
        const float bd = (o + R >= g.izbeg && o + R <= g.izend) ? sb[cur] : 0.0f;
        const AdjTiles<AX, AY>& t = S[par];
        // ---- A: this plane's loads before its stores (one counter tracks both): column words of plane o + 1, the
        // temporaries of plane p + 1 into the other buffer (last read before the previous barrier), tile of p + 2
        ncol[par ^ 1] = col_load(o + 8);
        if (p + 1 <= z1 + R - 1)
            commit(par ^ 1);
        if (p + 2 <= z1 + R - 1)
            fetch(p + 2);
        // ---- B: d_x, d_y of plane p
        // (the scale is zero for non-interior points: every tile value is finite)
#pragma unroll
        for (int i = 0; i < XI; ++i) {
            // items past the count (last slot only) repeat the last item: same value, same word
            if (i == 0 || tid + i * NT < FXI) {
                const float bb = t.Bt[xr[i] + 4][xj[i]] * (xs[i] * pin);
                float2 v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    v[k] = t.SX[xr[i]][xj[i] + k];
                X[xr[i]][xj[i]] = make_float2(bb * stg8([&](int k) { return v[k].x; }),
                                              bb * stg8([&](int k) { return v[k].y; }));
            }
        }
#pragma unroll
        for (int h = 0; h < YI; ++h) {
            const int j = ty + h * AY;
            if (h == 0 || j < BH) {
                const float bb = t.Bt[j][tx + 4] * (ys[h] * pin);
                float2 v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    v[k] = t.SY[j + k][tx];
                Y[j][tx] = make_float2(bb * stg8([&](int k) { return v[k].x; }), bb * stg8([&](int k) { return v[k].y; }));
            }
        }
        __syncthreads();
        // ---- C: plane-p contributions
        if (active) {
            float gx1 = 0, gx2 = 0, gy1 = 0, gy2 = 0, hxy1 = 0, hxy2 = 0;
#pragma unroll
            for (int n = 1; n <= R; ++n) {
                const float c = cstg(n);
                const float2 xp = X[ty][tx + 4 + n], xm = X[ty][tx + 5 - n];
                const float2 yp = Y[ty + 4 + n][tx], ym = Y[ty + 5 - n][tx];
                gx1 += c * (xp.x - xm.x);
                gx2 += c * (xp.y - xm.y);
                gy1 += c * (yp.x - ym.x);
                gy2 += c * (yp.y - ym.y);
                const float e = ecen(n);
                const float2 pp = t.MXY[ty + 4 + n][tx + 4 + n], pm = t.MXY[ty + 4 + n][tx + 4 - n];
                const float2 mm = t.MXY[ty + 4 - n][tx + 4 - n], mp = t.MXY[ty + 4 - n][tx + 4 + n];
                hxy1 += e * (pp.x - pm.x + mm.x - mp.x);
                hxy2 += e * (pp.y - pm.y + mm.y - mp.y);
                const float2 yzp = t.MYZ[ty + 4 + n][tx], yzm = t.MYZ[ty + 4 - n][tx];
                const float2 zxp = t.MZX[ty][tx + 4 + n], zxm = t.MZX[ty][tx + 4 - n];
                const float d1 = fmaf(i4yz, yzp.x - yzm.x, i4zx * (zxp.x - zxm.x));
                const float d2 = fmaf(i4yz, yzp.y - yzm.y, i4zx * (zxp.y - zxm.y));
                q1[cur + 4 - n] += e * d1;
                q1[cur + 4 + n] -= e * d1;
                q2[cur + 4 - n] += e * d2;
                q2[cur + 4 + n] -= e * d2;
            }
            // in-plane terms of output p (slot c + 4); explicit roundings: no product is fused into these sums
            q1[cur + 4] = __fadd_rn(q1[cur + 4], __fadd_rn(__fadd_rn(__fmul_rn(gx1, invdx), __fmul_rn(gy1, invdy)),
                                                           __fmul_rn(hxy1, i4xy)));
            q2[cur + 4] = __fadd_rn(q2[cur + 4], __fadd_rn(__fadd_rn(__fmul_rn(gx2, invdx), __fmul_rn(gy2, invdy)),
                                                           __fmul_rn(hxy2, i4xy)));
        }
        // ---- D: d_z(o+4), finalise plane o
        {
            const float dzs = bd * invdz;
            dz1[cur + 7] = dzs * stg8([&](int k) { return sz1[cur + k]; });
            dz2[cur + 7] = dzs * stg8([&](int k) { return sz2[cur + k]; });
            if (g.fs && o == g.izbeg) {
                mirror_dz<cur>(dz1);
                mirror_dz<cur>(dz2);
            }
        }
        const bool fin = o >= z0 && active;
        float un = 0.0f, wn = 0.0f, a1n = 0.0f, a2n = 0.0f;
        double gn = 0.0;
        const unsigned gis = gi[par];
        const float uo = pw[par][0], wo = pw[par][1], ao1 = sa1[cur], ao2 = sa2[cur];
        if (fin) {
            const float gz1 = stg8([&](int k) { return dz1[cur + k]; });
            const float gz2 = stg8([&](int k) { return dz2[cur + k]; });
            // fixed rounding whichever products the compiler can see (see the forward finalisation)
            const float r3 = fmaf(gz1, invdz, q1[cur]);
            const float r4 = fmaf(gz2, invdz, q2[cur]);
            float a = 1.0f, bb = 1.0f;
            if constexpr (SPONGE) {
                a = fminf(axy, abz[o].a);
                bb = fminf(bxy, abz[o].b);
            }
            un = SPONGE ? fmaf(bb * dt, r3, a * uo) : fmaf(r3, dt, uo);
            wn = SPONGE ? fmaf(bb * dt, r4, a * wo) : fmaf(r4, dt, wo);
            a1n = fmaf(bb * un, dt, a * ao1);
            a2n = fmaf(bb * wn, dt, a * ao2);
            if (imaging) {
                const float su = snap_decode(SNAP, sw_u[par]), sw = snap_decode(SNAP, sw_w[par]);
                gn = g0[par] + fma((double)wn, (double)sw, (double)un * (double)su);
            }
        }
        // pointwise words of plane o + 1 (their registers are free now), then the stores of plane o
        fetch_point(o + 1, par ^ 1);
        if (fin) {
            const unsigned io = (unsigned)o * PB + colb;
            stf(u, io, un);
            stf(w, io, wn);
            stf(a1o, io, a1n);
            stf(a2o, io, a2n);
            if (imaging)
                grad[gis] = gn;
        }
        // ---- E: barrier (the tile buffer and X/Y are rewritten by the next plane)
        __syncthreads();
        if (cur == AU - 1) {
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                sz1[k] = sz1[k + AU];
                sz2[k] = sz2[k + AU];
                dz1[k] = dz1[k + AU];
                dz2[k] = dz2[k + AU];
                sa1[k] = sa1[k + AU];
                sa2[k] = sa2[k + AU];
            }
#pragma unroll
            for (int k = 0; k < 3; ++k)
                sb[k] = sb[k + AU];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                q1[k] = q1[k + AU];
                q2[k] = q2[k + AU];
            }
#pragma unroll
            for (int k = 8; k < 8 + AU; ++k)
                q1[k] = q2[k] = 0.0f;
        }
    };

    for (int o = z0 - 8; o < z1; o += AU) {
        body(std::integral_constant<int, 0>{}, o);
        if (o + 1 >= z1)
            break;
        body(std::integral_constant<int, 1>{}, o + 1);
        if constexpr (AU == 4) {
            if (o + 2 >= z1)
                break;
            body(std::integral_constant<int, 2>{}, o + 2);
            if (o + 3 >= z1)
                break;
            body(std::integral_constant<int, 3>{}, o + 3);
        }
    }
}


// ------------------------------------------------------------------------------------------------------------------
// Receivers
// ------------------------------------------------------------------------------------------------------------------

// pressure -(2 s1 + s2)/3 at linear index i (simulator2/simulator2SetUp.cu:31-42)
__device__ __forceinline__ float
pressure_at(const float* __restrict__ s1, const float* __restrict__ s2, long i)
{
    return -(2.0f * s1[i] + s2[i]) / 3.0f;
}

// Transpose of the pressure recording at receiver k, applied after an adjoint step: u += -2/3 r, w += -1/3 r and the
// already-updated sigma_out gets B dt times the same increment (the adjoint step's sigma' = A sigma + B dt u').
// Optionally adds the increment's share of the imaging condition.
__device__ __forceinline__ void
inject_point(float* __restrict__ u, float* __restrict__ w, float* __restrict__ a1o, float* __restrict__ a2o,
             const long* __restrict__ idx, const int* __restrict__ ijk, int k, const float* __restrict__ res, float dt,
             const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, const Box& g,
             const Snap& snap, double* __restrict__ grad)
{
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
        const long sidx = row * (snap.ldx / SNAP_SEG) + (x - g.ixbeg) / SNAP_SEG;
        const long si = row * snap.ldx + (x - g.ixbeg);
        const float su = snap_load(snap.mode, snap.a, snap.sa, si, sidx);
        const float sw = snap_load(snap.mode, snap.b, snap.sb, si, sidx);
        const long gstride_y = (long)(g.ixend - g.ixbeg + 1), gstride_z = gstride_y * (g.iyend - g.iybeg + 1);
        const long gi = (long)(z - g.izbeg) * gstride_z + (long)(y - g.iybeg) * gstride_y + (x - g.ixbeg);
        grad[gi] += fma((double)dw, (double)sw, (double)du * (double)su);
    }
}


// ------------------------------------------------------------------------------------------------------------------
// Host-side argument checks shared by the launchers
// ------------------------------------------------------------------------------------------------------------------

// The fused kernels stage tiles with 16-byte loads: the interior must start on a 4-float boundary, rows must be a
// multiple of 4 floats and every volume 16-byte aligned.
inline bool
box_ok(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy)
{
    return ixbeg >= R && iybeg >= R && izbeg >= R && ixend >= ixbeg && iyend >= iybeg && izend >= izbeg &&
           ixend + R < ldimx && iyend + R < ldimy && ixbeg % 4 == 0 && ldimx % 4 == 0;
}

// Every volume (allocated planes up to izend + 4) must be smaller than 4 GiB: the kernels address it with 32-bit
// byte offsets.
inline bool
small_volume(int ldimx, int ldimy, int izend)
{
    return (long)ldimx * ldimy * (izend + R + 1) * (long)sizeof(float) < (1L << 32);
}

inline bool
aligned16(std::initializer_list<const void*> ps)
{
    for (const void* p : ps)
        if (reinterpret_cast<uintptr_t>(p) % 16 != 0)
            return false;
    return true;
}

inline bool
snap_ok(const tti_snapshot* s, int nx, int ny, int nz)
{
    if (s == nullptr || s->mode == TTI_SNAP_NONE)
        return true;
    if (s->mode < TTI_SNAP_FP32 || s->mode > TTI_SNAP_BF16 || !s->field3 || !s->field4)
        return false;
    if (s->ldimx % 64 != 0 || s->ldimx < nx || s->ldimy < ny)
        return false;
    if (s->mode == TTI_SNAP_FP16 && (!s->scale3 || !s->scale4))
        return false;
    if ((s->field1 == nullptr) != (s->field2 == nullptr))
        return false;
    if (s->field1 && s->mode == TTI_SNAP_FP16 && (!s->scale1 || !s->scale2))
        return false;
    // the kernels address a snapshot field with 32-bit element offsets
    return (long)nz * s->ldimy * s->ldimx < (1L << 32);
}

inline Snap
to_snap(const tti_snapshot* s)
{
    Snap r{ TTI_SNAP_NONE, nullptr, nullptr, nullptr, nullptr, 64, 1, nullptr, nullptr, nullptr, nullptr };
    if (s && s->mode != TTI_SNAP_NONE) {
        r.mode = s->mode;
        r.a = s->field3;
        r.b = s->field4;
        r.sa = s->scale3;
        r.sb = s->scale4;
        r.ldx = s->ldimx;
        r.ldy = s->ldimy;
        r.c = s->field1;
        r.d = s->field2;
        r.sc = s->scale1;
        r.sd = s->scale2;
    }
    return r;
}


} // namespace fused
} // namespace tti
