// Fixed-rate ZFP codec of 3-D FP32 fields for gfx950: the compressed-snapshot path of the snapshot store.
//
// The reference hands every snapshot volume to the zfp library (LLNL) in fixed-rate mode on the GPU
// (snap_helpers/rtm_snap_gpu.cu:228-280, set up at simulator2/simulator2SetUp.cu:285-323). The stream format is that
// library's fixed-rate layout, restated from its published algorithm (P. Lindstrom, "Fixed-Rate Compressed
// Floating-Point Arrays", IEEE TVCG 2014):
//   * 4x4x4 blocks in raster order (x fastest), each exactly maxbits = floor(64 rate + 0.5) bits, LSB first in 64-bit
//     words; partial blocks at the far faces are padded per axis (n = 1: a a a a, 2: a b b a, 3: a b c a);
//   * per block: a "nonzero" bit, 8 bits of the biased common exponent (frexp exponent of the largest magnitude, at
//     least -126), then the 64 values scaled by 2^(30 - emax) and truncated to int32, decorrelated by the lifting
//     transform along x, y, z, put in sequency order, mapped to negabinary and coded bit plane by bit plane (the
//     first n bits of a plane verbatim, the rest by group tests and unary runs), then zero padding.
//
// GPU mapping: one thread codes one block (the coder is a sequential bit-serial walk whose length depends on the
// data; 64 blocks per wave run it in lock-step). Thread t of a wave takes the t-th block along x, so the 16 row
// loads of a block are 16-byte pieces of 1 KB contiguous wave accesses (float4 loads when the field is 16-byte
// aligned with row and plane strides that are multiples of 4). The 64 coefficients stay in registers: every loop
// over them is unrolled so each array index is a constant. A bit plane is gathered with 64 bit-field extracts; the
// unary runs are written and parsed a run at a time (count trailing zeros), not a bit at a time. Blocks whose size
// is a whole number of words are written with plain stores; otherwise the word a block shares with its neighbour is
// merged with a 64-bit atomic OR into a stream cleared beforehand.
//
// Scaling to integers uses the float's bits (exact for every value, FTZ or not); the library multiplies by the float
// 2^(30 - emax), which is infinite for blocks whose largest magnitude is below 2^-97, the only blocks coded
// differently.
#include <cstdint>

#include "common.h"
#include "tti_zfp.h"

namespace tti {
namespace zfp {

using u64 = unsigned long long;

constexpr unsigned NBMASK = 0xaaaaaaaau;
constexpr int NT = 256; // threads per workgroup

// Sequency order: the coefficient (i, j, k) at position i + 4j + 16k, sorted by i + j + k, then i^2 + j^2 + k^2.
__host__ __device__ constexpr int
perm3(int p)
{
#define ZI(i, j, k) ((i) + 4 * (j) + 16 * (k))
    constexpr unsigned char t[64] = {
        ZI(0, 0, 0), ZI(1, 0, 0), ZI(0, 1, 0), ZI(0, 0, 1), ZI(0, 1, 1), ZI(1, 0, 1), ZI(1, 1, 0), ZI(2, 0, 0),
        ZI(0, 2, 0), ZI(0, 0, 2), ZI(1, 1, 1), ZI(2, 1, 0), ZI(2, 0, 1), ZI(0, 2, 1), ZI(1, 2, 0), ZI(1, 0, 2),
        ZI(0, 1, 2), ZI(3, 0, 0), ZI(0, 3, 0), ZI(0, 0, 3), ZI(2, 1, 1), ZI(1, 2, 1), ZI(1, 1, 2), ZI(0, 2, 2),
        ZI(2, 0, 2), ZI(2, 2, 0), ZI(3, 1, 0), ZI(3, 0, 1), ZI(0, 3, 1), ZI(1, 3, 0), ZI(1, 0, 3), ZI(0, 1, 3),
        ZI(1, 2, 2), ZI(2, 1, 2), ZI(2, 2, 1), ZI(3, 1, 1), ZI(1, 3, 1), ZI(1, 1, 3), ZI(3, 2, 0), ZI(3, 0, 2),
        ZI(0, 3, 2), ZI(2, 3, 0), ZI(2, 0, 3), ZI(0, 2, 3), ZI(2, 2, 2), ZI(3, 2, 1), ZI(3, 1, 2), ZI(1, 3, 2),
        ZI(2, 3, 1), ZI(2, 1, 3), ZI(1, 2, 3), ZI(0, 3, 3), ZI(3, 0, 3), ZI(3, 3, 0), ZI(3, 2, 2), ZI(2, 3, 2),
        ZI(2, 2, 3), ZI(1, 3, 3), ZI(3, 1, 3), ZI(3, 3, 1), ZI(2, 3, 3), ZI(3, 2, 3), ZI(3, 3, 2), ZI(3, 3, 3),
    };
#undef ZI
    return t[p];
}

// Lifting transform of four values: 1/16 [4 4 4 4; 5 1 -1 -5; -4 4 4 -4; -2 6 -6 2] with the rounding of the
// arithmetic right shifts, and its inverse.
__device__ __forceinline__ void
fwd_lift(int& x, int& y, int& z, int& w)
{
    x += w; x >>= 1; w -= x;
    z += y; z >>= 1; y -= z;
    x += z; x >>= 1; z -= x;
    w += y; w >>= 1; y -= w;
    w += y >> 1; y -= w >> 1;
}

__device__ __forceinline__ void
inv_lift(int& x, int& y, int& z, int& w)
{
    y += w >> 1; w -= y >> 1;
    y += w; w *= 2; w -= y;
    z += x; x *= 2; x -= z;
    y += z; z *= 2; z -= y;
    w += x; x *= 2; x -= w;
}

__device__ __forceinline__ void
fwd_xform(int (&q)[64])
{
#pragma unroll
    for (int r = 0; r < 16; ++r) { // along x: rows (y, z)
        const int o = 4 * r;
        fwd_lift(q[o], q[o + 1], q[o + 2], q[o + 3]);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) { // along y: (x, z)
        const int o = (r & 3) + 16 * (r >> 2);
        fwd_lift(q[o], q[o + 4], q[o + 8], q[o + 12]);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) // along z: (x, y)
        fwd_lift(q[r], q[r + 16], q[r + 32], q[r + 48]);
}

__device__ __forceinline__ void
inv_xform(int (&q)[64])
{
#pragma unroll
    for (int r = 0; r < 16; ++r)
        inv_lift(q[r], q[r + 16], q[r + 32], q[r + 48]);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int o = (r & 3) + 16 * (r >> 2);
        inv_lift(q[o], q[o + 4], q[o + 8], q[o + 12]);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int o = 4 * r;
        inv_lift(q[o], q[o + 1], q[o + 2], q[o + 3]);
    }
}

// trunc(x 2^(30 - emax)) from the float's bits; |x| < 2^emax keeps the result below 2^30 in magnitude
__device__ __forceinline__ int
to_fixed(float x, int emax)
{
    const unsigned u = __float_as_uint(x);
    const int e = (int)((u >> 23) & 0xffu);
    const unsigned mant = e ? ((u & 0x7fffffu) | 0x800000u) : (u & 0x7fffffu);
    const int sh = max(e, 1) - 120 - emax; // value = mant 2^(max(e,1) - 150)
    unsigned m;
    if (sh >= 0)
        m = mant << sh;
    else
        m = -sh < 32 ? mant >> -sh : 0u;
    return (u >> 31) ? -(int)m : (int)m;
}

// source offset along one axis of a block with n valid values (zfp's padding of partial blocks)
__device__ __forceinline__ int
pad_at(int i, int n)
{
    if (n >= 4)
        return i;
    if (n == 1)
        return 0;
    if (n == 2)
        return (i == 1 || i == 2) ? 1 : 0;
    return i == 3 ? 0 : i;
}

struct BitWriter
{
    u64* w;
    long wi;   // index of the word being filled
    u64 acc;   // its bits so far
    int fill;  // bits of acc in use (0..63)
    bool head; // the word is shared with the previous block (merge it)

    __device__ BitWriter(u64* words, u64 off)
        : w(words), wi((long)(off >> 6)), acc(0), fill((int)(off & 63)), head((off & 63) != 0)
    {
    }
    __device__ __forceinline__ void emit()
    {
        if (head) {
            atomicOr(&w[wi], acc);
            head = false;
        } else {
            w[wi] = acc;
        }
        ++wi;
    }
    // the n (0..64) low bits of v; v holds no higher bits
    __device__ __forceinline__ void put(u64 v, int n)
    {
        if (n <= 0)
            return;
        acc |= v << fill;
        if (fill + n >= 64) {
            emit();
            acc = fill ? v >> (64 - fill) : 0ull;
            fill += n - 64;
        } else {
            fill += n;
        }
    }
    __device__ void pad(unsigned n)
    {
        while (n) {
            const int c = n < 64u ? (int)n : 64;
            put(0ull, c);
            n -= (unsigned)c;
        }
    }
    // the last word is shared with the next block when the block ends inside it
    __device__ void finish()
    {
        if (fill)
            atomicOr(&w[wi], acc);
    }
};

struct BitReader
{
    const u64* w;
    long wi;    // next word to load
    u64 buf;    // unread bits of the current word, next bit at the LSB
    int avail;  // how many

    __device__ BitReader(const u64* words, u64 off) : w(words), wi((long)(off >> 6))
    {
        const int s = (int)(off & 63);
        buf = w[wi++] >> s;
        avail = 64 - s;
    }
    // n in 0..64 bits; never loads a word beyond the last bit it returns
    __device__ __forceinline__ u64 get(int n)
    {
        if (n <= 0)
            return 0ull;
        if (n <= avail) {
            const u64 r = n == 64 ? buf : buf & ((1ull << n) - 1);
            buf = n == 64 ? 0ull : buf >> n;
            avail -= n;
            return r;
        }
        const u64 nw = w[wi++];
        const int need = n - avail; // 1..64
        u64 r = avail ? (buf | (nw << avail)) : nw;
        r = n == 64 ? r : r & ((1ull << n) - 1);
        buf = need == 64 ? 0ull : nw >> need;
        avail = 64 - need;
        return r;
    }
    // the next n (0..63) bits without consuming them
    __device__ __forceinline__ u64 peek(int n) const
    {
        if (n <= 0)
            return 0ull;
        const u64 mask = (1ull << n) - 1;
        if (n <= avail)
            return buf & mask;
        const u64 nw = w[wi];
        return (avail ? (buf | (nw << avail)) : nw) & mask;
    }
};

struct Geom
{
    int nx, ny, nz, nbx, nby;
    long sy, sz;
    long nblocks;
};

__device__ __forceinline__ void
block_coords(const Geom& g, long b, int& x0, int& y0, int& z0)
{
    const long bxy = (long)g.nbx * g.nby;
    const int bz = (int)(b / bxy);
    const long r = b - (long)bz * bxy;
    const int by = (int)(r / g.nbx);
    const int bx = (int)(r - (long)by * g.nbx);
    x0 = 4 * bx;
    y0 = 4 * by;
    z0 = 4 * bz;
}

template <bool VEC>
__global__ __launch_bounds__(NT) void
k_zfp_encode(const float* __restrict__ f, Geom g, unsigned maxbits, u64* __restrict__ words)
{
    const long b = (long)blockIdx.x * NT + threadIdx.x;
    if (b >= g.nblocks)
        return;
    int x0, y0, z0;
    block_coords(g, b, x0, y0, z0);
    const int mx = g.nx - x0, my = g.ny - y0, mz = g.nz - z0;
    float v[64];
    if (VEC && mx >= 4 && my >= 4 && mz >= 4) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float4 q = *reinterpret_cast<const float4*>(f + (long)(z0 + (r >> 2)) * g.sz +
                                                              (long)(y0 + (r & 3)) * g.sy + x0);
            v[4 * r] = q.x;
            v[4 * r + 1] = q.y;
            v[4 * r + 2] = q.z;
            v[4 * r + 3] = q.w;
        }
    } else {
        int px[4], py[4], pz[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            px[i] = x0 + pad_at(i, mx);
            py[i] = y0 + pad_at(i, my);
            pz[i] = z0 + pad_at(i, mz);
        }
#pragma unroll
        for (int i = 0; i < 64; ++i)
            v[i] = f[(long)pz[i >> 4] * g.sz + (long)py[(i >> 2) & 3] * g.sy + px[i & 3]];
    }

    BitWriter out(words, (u64)b * maxbits);
    unsigned mbits = 0;
#pragma unroll
    for (int i = 0; i < 64; ++i)
        mbits = max(mbits, __float_as_uint(v[i]) & 0x7fffffffu);
    if (!mbits) { // all zero: one zero bit, then padding
        out.pad(maxbits);
        out.finish();
        return;
    }
    const int ebits = (int)(mbits >> 23);
    const int emax = ebits ? ebits - 126 : -126;
    out.put(2ull * (unsigned)(emax + 127) + 1ull, 9);

    int q[64];
#pragma unroll
    for (int i = 0; i < 64; ++i)
        q[i] = to_fixed(v[i], emax);
    fwd_xform(q);
    unsigned u[64];
#pragma unroll
    for (int i = 0; i < 64; ++i)
        u[i] = ((unsigned)q[perm3(i)] + NBMASK) ^ NBMASK;

    unsigned bits = maxbits - 9, n = 0;
    for (int k = 31; k >= 0 && bits; --k) {
        u64 x = 0;
#pragma unroll
        for (int i = 0; i < 64; ++i)
            x |= (u64)((u[i] >> k) & 1u) << i;
        // the first n bits verbatim
        const unsigned m = min(n, bits);
        bits -= m;
        out.put(m == 64 ? x : x & ((1ull << m) - 1), (int)m);
        x = m == 64 ? 0ull : x >> m;
        // group test, zeros up to the next one bit, that bit (implicit at position 63)
        while (n < 64 && bits) {
            --bits;
            out.put(x ? 1ull : 0ull, 1);
            if (!x)
                break;
            const unsigned tz = (unsigned)__builtin_ctzll(x), room = 63 - n;
            if (tz < room) {
                if (tz + 1 <= bits) {
                    out.put(1ull << tz, (int)tz + 1);
                    bits -= tz + 1;
                    n += tz;
                    x >>= tz;
                } else {
                    out.put(0ull, (int)bits);
                    bits = 0;
                }
            } else {
                if (room <= bits) {
                    out.put(0ull, (int)room);
                    bits -= room;
                    n += room;
                    x >>= room;
                } else {
                    out.put(0ull, (int)bits);
                    bits = 0;
                }
            }
            x >>= 1;
            ++n;
        }
    }
    out.pad(bits);
    out.finish();
}

template <bool VEC>
__global__ __launch_bounds__(NT) void
k_zfp_decode(const u64* __restrict__ words, Geom g, unsigned maxbits, float* __restrict__ f)
{
    const long b = (long)blockIdx.x * NT + threadIdx.x;
    if (b >= g.nblocks)
        return;
    int x0, y0, z0;
    block_coords(g, b, x0, y0, z0);
    float v[64];
    BitReader in(words, (u64)b * maxbits);
    if (!in.get(1)) {
#pragma unroll
        for (int i = 0; i < 64; ++i)
            v[i] = 0.0f;
    } else {
        const int emax = (int)in.get(8) - 127;
        unsigned u[64];
#pragma unroll
        for (int i = 0; i < 64; ++i)
            u[i] = 0u;
        unsigned bits = maxbits - 9, n = 0;
        for (int k = 31; k >= 0 && bits; --k) {
            const unsigned m = min(n, bits);
            bits -= m;
            u64 x = in.get((int)m);
            while (n < 64 && bits) {
                --bits;
                if (!in.get(1))
                    break;
                const unsigned lim = min(63 - n, bits);
                const u64 p = in.peek((int)lim);
                if (p) {
                    const unsigned t = (unsigned)__builtin_ctzll(p);
                    in.get((int)t + 1);
                    bits -= t + 1;
                    n += t;
                } else {
                    in.get((int)lim);
                    bits -= lim;
                    n += lim;
                }
                x += 1ull << n;
                ++n;
            }
#pragma unroll
            for (int i = 0; i < 64; ++i)
                u[i] |= (unsigned)((x >> i) & 1ull) << k;
        }
        int q[64];
#pragma unroll
        for (int i = 0; i < 64; ++i)
            q[perm3(i)] = (int)((u[i] ^ NBMASK) - NBMASK);
        inv_xform(q);
#pragma unroll
        for (int i = 0; i < 64; ++i)
            v[i] = ldexpf((float)q[i], emax - 30);
    }
    const int mx = g.nx - x0, my = g.ny - y0, mz = g.nz - z0;
    if (VEC && mx >= 4 && my >= 4 && mz >= 4) {
#pragma unroll
        for (int r = 0; r < 16; ++r)
            *reinterpret_cast<float4*>(f + (long)(z0 + (r >> 2)) * g.sz + (long)(y0 + (r & 3)) * g.sy + x0) =
                make_float4(v[4 * r], v[4 * r + 1], v[4 * r + 2], v[4 * r + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < 64; ++i)
            if ((i & 3) < mx && ((i >> 2) & 3) < my && (i >> 4) < mz)
                f[(long)(z0 + (i >> 4)) * g.sz + (long)(y0 + ((i >> 2) & 3)) * g.sy + x0 + (i & 3)] = v[i];
    }
}

__global__ void
k_zero_words(u64* __restrict__ w, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        w[i] = 0ull;
}

bool
geom(int nx, int ny, int nz, long sy, long sz, double rate, Geom& g)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || sy < nx || sz < sy * (long)ny || !(rate > 0.0) || rate > 64.0)
        return false;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.sy = sy;
    g.sz = sz;
    g.nbx = (nx + 3) / 4;
    g.nby = (ny + 3) / 4;
    g.nblocks = (long)g.nbx * g.nby * ((nz + 3) / 4);
    return true;
}

bool
vec_ok(const void* p, long sy, long sz)
{
    return (reinterpret_cast<uintptr_t>(p) & 15u) == 0 && sy % 4 == 0 && sz % 4 == 0;
}

size_t
words_of(long nblocks, unsigned maxbits)
{
    return (size_t)(((u64)nblocks * maxbits + 63) / 64);
}

} // namespace zfp
} // namespace tti

using namespace tti;
using namespace tti::zfp;

extern "C" {

unsigned
tti_zfp_block_bits(double rate)
{
    if (!(rate > 0.0))
        return 9u;
    const double b = __builtin_floor(64.0 * rate + 0.5);
    return b < 9.0 ? 9u : (unsigned)b;
}

size_t
tti_zfp_stream_bytes(double rate, int nx, int ny, int nz)
{
    Geom g;
    if (!geom(nx, ny, nz, nx, (long)nx * ny, rate, g))
        return 0;
    return 8 * words_of(g.nblocks, tti_zfp_block_bits(rate));
}

size_t
tti_zfp_maximum_size(double rate, int nx, int ny, int nz)
{
    Geom g;
    if (!geom(nx, ny, nz, nx, (long)nx * ny, rate, g))
        return 0;
    return 8 * (size_t)((148ull + (u64)g.nblocks * tti_zfp_block_bits(rate) + 63) / 64);
}

int
tti_zfp_compress_3d(const float* field, int nx, int ny, int nz, long sy, long sz, double rate, void* out,
                    size_t out_bytes, size_t* compressed_bytes, void* stream)
{
    Geom g;
    if (!field || !out || !geom(nx, ny, nz, sy, sz, rate, g) || (reinterpret_cast<uintptr_t>(out) & 7u))
        return TTI_EINVAL;
    const unsigned maxbits = tti_zfp_block_bits(rate);
    const size_t nw = words_of(g.nblocks, maxbits);
    if (out_bytes < 8 * nw)
        return TTI_EINVAL;
    const hipStream_t s = (hipStream_t)stream;
    u64* w = static_cast<u64*>(out);
    if (maxbits % 64) // blocks share words: merged with atomic OR into a cleared stream
        k_zero_words<<<(unsigned)((nw + 255) / 256), 256, 0, s>>>(w, (long)nw);
    const unsigned grid = (unsigned)((g.nblocks + NT - 1) / NT);
    if (vec_ok(field, sy, sz))
        k_zfp_encode<true><<<grid, NT, 0, s>>>(field, g, maxbits, w);
    else
        k_zfp_encode<false><<<grid, NT, 0, s>>>(field, g, maxbits, w);
    if (compressed_bytes)
        *compressed_bytes = 8 * nw;
    return launch_status();
}

int
tti_zfp_decompress_3d(const void* in, size_t in_bytes, float* field, int nx, int ny, int nz, long sy, long sz,
                      double rate, void* stream)
{
    Geom g;
    if (!field || !in || !geom(nx, ny, nz, sy, sz, rate, g) || (reinterpret_cast<uintptr_t>(in) & 7u))
        return TTI_EINVAL;
    const unsigned maxbits = tti_zfp_block_bits(rate);
    if (in_bytes < 8 * words_of(g.nblocks, maxbits))
        return TTI_EINVAL;
    const hipStream_t s = (hipStream_t)stream;
    const unsigned grid = (unsigned)((g.nblocks + NT - 1) / NT);
    if (vec_ok(field, sy, sz))
        k_zfp_decode<true><<<grid, NT, 0, s>>>(static_cast<const u64*>(in), g, maxbits, field);
    else
        k_zfp_decode<false><<<grid, NT, 0, s>>>(static_cast<const u64*>(in), g, maxbits, field);
    return launch_status();
}

} // extern "C"
