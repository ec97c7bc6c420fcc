// The type2 and type1 operators at the reference's other FD orders: 16 for every launcher, and 12 for the density
// gradient (kernels/kernel_type2_fwd_drv.cu:381-395, kernels/kernel_type2_fwd.cu:547, :648-657,
// kernels/kernel_type2_grad_snap.cu:445-475, kernels/kernel_type2_adj.cu:295-335, kernels/kernel_type2_grad.cu:612-623,
// kernels/kernel_type1_gpu_fwd.cu:464-530, kernels/kernel_type1_gpu_adj.cu:420-480). Radius RO = order / 2.
//
// One thread per interior point, operands read straight from HBM through the caches: the order-8 launchers keep their
// LDS-tiled 2.5-D kernels; these serve the less used orders with the same arithmetic (operator definitions as in the
// order-8 kernels) and no scratch memory. The adjoint and density-gradient kernels form the pointwise temporaries
// S_{f,a}, M_{f,ab} (or the type1 T) at each stencil point they read instead of staging them.
//
// Coefficients (_FD_COEF_LEAST_SQUARE, FP32 as the reference spells them):
//   order 16: first staggered c1..c8 kernels/fd_coef.h:479-486, central second e0..e8 :535-543
//   order 12: first staggered c1..c6 of stg_d1<12> (kernels/fd_coefficients.h:16-60); no central set (density
//             gradient only)
//   order 8 : as common.h (used by the order-8 instantiations, which the parity tests compare with the tiled kernels)
#include "common.h"
#include "order_kernels.h"

#include <type_traits>

namespace tti {
namespace order {

template <int RO>
struct FD;
template <>
struct FD<4>
{
    static constexpr float c[5] = { 0.0f, C1, C2, C3, C4 };
    static constexpr float e[5] = { E0, E1, E2, E3, E4 };
};
template <>
struct FD<6>
{
    static constexpr float c[7] = { 0.0f, 1.2403120e+0f, -1.1141390e-1f, 2.5759060e-2f, -6.4207590e-3f,
                                    1.3095870e-3f, -1.5512770e-4f };
    static constexpr float e[7] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
};
template <>
struct FD<8>
{
    static constexpr float c[9] = { 0.0f, 1.254298E+00f, -1.234902E-01f, 3.470321E-02f, -1.201822E-02f,
                                    4.188332E-03f, -1.310232E-03f, 3.233353E-04f, -4.675068E-05f };
    static constexpr float e[9] = { -0.3162230E+1f, 0.1875943E+1f, -0.3859840E+0f, 0.1228261E+0f, -0.4233006E-1f,
                                    0.1379639E-1f, -0.3846673E-2f, 0.8026637E-3f, -0.9257725E-4f };
};

struct Geo
{
    int ixbeg, ixend, iybeg, iyend, izbeg, izend;
    long ldx, ldy;
};

// interior point of this thread (grid: x blocks of 64, y = rows, z = planes), false beyond ixend
__device__ __forceinline__ bool
point(const Geo& g, int& ix, int& iy, int& iz, long& i)
{
    ix = g.ixbeg + blockIdx.x * 64 + threadIdx.x;
    iy = g.iybeg + blockIdx.y;
    iz = g.izbeg + blockIdx.z;
    i = ((long)iz * g.ldy + iy) * g.ldx + ix;
    return ix <= g.ixend;
}

// D- g(i) = invd sum_n c_n [g(i+n-1) - g(i-n)] (value at i-1/2); D+ g(i) = invd sum_n c_n [g(i+n) - g(i-n+1)]
template <int RO, class F>
__device__ __forceinline__ float
dminus(F&& g, long i, long s)
{
    float a = 0.0f;
#pragma unroll
    for (int n = 1; n <= RO; ++n)
        a += FD<RO>::c[n] * (g(i + (n - 1) * s) - g(i - n * s));
    return a;
}
template <int RO, class F>
__device__ __forceinline__ float
dplus(F&& g, long i, long s)
{
    float a = 0.0f;
#pragma unroll
    for (int n = 1; n <= RO; ++n)
        a += FD<RO>::c[n] * (g(i + n * s) - g(i - (n - 1) * s));
    return a;
}
// X_ab g = sum_n e_n [(g(+n,+n) - g(+n,-n)) + (g(-n,-n) - g(-n,+n))] (times 1/(4 da db) by the caller)
template <int RO, class F>
__device__ __forceinline__ float
mixed(F&& g, long i, long sa, long sb)
{
    float a = 0.0f;
#pragma unroll
    for (int n = 1; n <= RO; ++n)
        a += FD<RO>::e[n] * ((g(i + n * sa + n * sb) - g(i + n * sa - n * sb)) +
                             (g(i - n * sa - n * sb) - g(i - n * sa + n * sb)));
    return a;
}
// central second derivative e0 g + sum_n e_n (g(i-n) + g(i+n)) (times 1/da^2 by the caller)
template <int RO, class F>
__device__ __forceinline__ float
central2(F&& g, long i, long s)
{
    float a = FD<RO>::e[0] * g(i);
#pragma unroll
    for (int n = 1; n <= RO; ++n)
        a += FD<RO>::e[n] * (g(i - n * s) + g(i + n * s));
    return a;
}

// TTI combination at one point (the F_aa arrive scaled by rho where the operator has it)
__device__ __forceinline__ void
combine(float m1, float m2, float m3, float m4, float rx, float ry, float rz, const float (&f1)[6],
        const float (&f2)[6], float& r3, float& r4)
{
    // f[] = xx yy zz xy yz zx
    const float rx2 = rx * rx, ry2 = ry * ry, rz2 = rz * rz;
    const float rxy = 2.0f * rx * ry, ryz = 2.0f * ry * rz, rzx = 2.0f * rz * rx;
    const float v1 = rx2 * f1[0] + ry2 * f1[1] + rz2 * f1[2] + rxy * f1[3] + ryz * f1[4] + rzx * f1[5];
    const float h1 = f1[0] + f1[1] + f1[2] - v1;
    const float v2 = rx2 * f2[0] + ry2 * f2[1] + rz2 * f2[2] + rxy * f2[3] + ryz * f2[4] + rzx * f2[5];
    const float h2 = f2[0] + f2[1] + f2[2] - v2;
    r3 = m1 * h1 + m4 * v1 + m2 * v2;
    r4 = m2 * h1 + m3 * v2 + m4 * h2;
}

// pointwise adjoint temporaries of a sigma pair: lap, hz of each field (kernels/kernel_type2_grad_snap.cu:66-135)
struct Lh
{
    float lap1, hz1, lap2, hz2;
};
__device__ __forceinline__ Lh
lap_hz(float f1, float f2, float m1, float m2, float m3, float m4)
{
    return Lh{ m1 * f1 + m2 * f2, (-m1 + m4) * f1 - m2 * f2, m4 * f2, m2 * f1 + (-m4 + m3) * f2 };
}

struct Vols
{
    const float *f1, *f2, *m1, *m2, *m3, *m4, *rx, *ry, *rz, *b;
    __device__ __forceinline__ Lh lh(long j) const { return lap_hz(f1[j], f2[j], m1[j], m2[j], m3[j], m4[j]); }
    __device__ __forceinline__ float rho(long j) const { return 1.0f / b[j]; }
};

// ---- a2 forward derivatives d = b D- f -----------------------------------------------------------------------------
template <int RO>
__global__ __launch_bounds__(64) void k_fwd_drv(const float* __restrict__ f1, const float* __restrict__ f2,
                                                const float* __restrict__ b, float* __restrict__ d1x,
                                                float* __restrict__ d1y, float* __restrict__ d1z,
                                                float* __restrict__ d2x, float* __restrict__ d2y,
                                                float* __restrict__ d2z, Geo g, float ix_, float iy_, float iz_)
{
    int ix, iy, iz;
    long i;
    if (!point(g, ix, iy, iz, i))
        return;
    const long sy = g.ldx, sz = g.ldx * g.ldy;
    auto F1 = [&](long j) { return f1[j]; };
    auto F2 = [&](long j) { return f2[j]; };
    const float bb = b[i];
    d1x[i] = dminus<RO>(F1, i, 1) * ix_ * bb;
    d1y[i] = dminus<RO>(F1, i, sy) * iy_ * bb;
    d1z[i] = dminus<RO>(F1, i, sz) * iz_ * bb;
    d2x[i] = dminus<RO>(F2, i, 1) * ix_ * bb;
    d2y[i] = dminus<RO>(F2, i, sy) * iy_ * bb;
    d2z[i] = dminus<RO>(F2, i, sz) * iz_ * bb;
}

// ---- a3 forward update u += dt rhs ----------------------------------------------------------------------------------
template <int RO>
__global__ __launch_bounds__(64) void k_fwd_update(float* __restrict__ u, float* __restrict__ w,
                                                   const float* __restrict__ d1x, const float* __restrict__ d2x,
                                                   const float* __restrict__ d1y, const float* __restrict__ d2y,
                                                   const float* __restrict__ d1z, const float* __restrict__ d2z,
                                                   Vols v, Geo g, float dt, float ix_, float iy_, float iz_,
                                                   float ixy, float iyz, float izx)
{
    int ix, iy, iz;
    long i;
    if (!point(g, ix, iy, iz, i))
        return;
    const long sy = g.ldx, sz = g.ldx * g.ldy;
    auto G = [](const float* p) { return [p](long j) { return p[j]; }; };
    auto F1 = G(v.f1);
    auto F2 = G(v.f2);
    const float rho = v.rho(i);
    const float a1[6] = { rho * (dplus<RO>(G(d1x), i, 1) * ix_), rho * (dplus<RO>(G(d1y), i, sy) * iy_),
                          rho * (dplus<RO>(G(d1z), i, sz) * iz_), mixed<RO>(F1, i, 1, sy) * ixy,
                          mixed<RO>(F1, i, sy, sz) * iyz, mixed<RO>(F1, i, sz, 1) * izx };
    const float a2[6] = { rho * (dplus<RO>(G(d2x), i, 1) * ix_), rho * (dplus<RO>(G(d2y), i, sy) * iy_),
                          rho * (dplus<RO>(G(d2z), i, sz) * iz_), mixed<RO>(F2, i, 1, sy) * ixy,
                          mixed<RO>(F2, i, sy, sz) * iyz, mixed<RO>(F2, i, sz, 1) * izx };
    float r3, r4;
    combine(v.m1[i], v.m2[i], v.m3[i], v.m4[i], v.rx[i], v.ry[i], v.rz[i], a1, a2, r3, r4);
    u[i] += r3 * dt;
    w[i] += r4 * dt;
}

// ---- a6 adjoint derivatives d = b D- S, rhs = sum X_ab M ------------------------------------------------------------
template <int RO>
__global__ __launch_bounds__(64) void k_adj_drv(float* __restrict__ d1x, float* __restrict__ d2x,
                                                float* __restrict__ d1y, float* __restrict__ d2y,
                                                float* __restrict__ d1z, float* __restrict__ d2z,
                                                float* __restrict__ r3, float* __restrict__ r4, Vols v, Geo g,
                                                float ix_, float iy_, float iz_, float ixy, float iyz, float izx)
{
    int ix, iy, iz;
    long i;
    if (!point(g, ix, iy, iz, i))
        return;
    const long sy = g.ldx, sz = g.ldx * g.ldy;
    // S_{f,a}(j) = rho (lap_f + r_a^2 hz_f), M_{f,ab}(j) = 2 r_a r_b hz_f, formed at every stencil point
    auto S = [&](int f, const float* r) {
        return [&v, f, r](long j) {
            const Lh t = v.lh(j);
            const float ra = r[j];
            return v.rho(j) * (f == 1 ? t.lap1 + ra * ra * t.hz1 : t.lap2 + ra * ra * t.hz2);
        };
    };
    auto M = [&](int f, const float* ra, const float* rb) {
        return [&v, f, ra, rb](long j) {
            const Lh t = v.lh(j);
            return 2.0f * ra[j] * rb[j] * (f == 1 ? t.hz1 : t.hz2);
        };
    };
    const float bb = v.b[i];
    d1x[i] = dminus<RO>(S(1, v.rx), i, 1) * ix_ * bb;
    d2x[i] = dminus<RO>(S(2, v.rx), i, 1) * ix_ * bb;
    d1y[i] = dminus<RO>(S(1, v.ry), i, sy) * iy_ * bb;
    d2y[i] = dminus<RO>(S(2, v.ry), i, sy) * iy_ * bb;
    d1z[i] = dminus<RO>(S(1, v.rz), i, sz) * iz_ * bb;
    d2z[i] = dminus<RO>(S(2, v.rz), i, sz) * iz_ * bb;
    r3[i] = mixed<RO>(M(1, v.rx, v.ry), i, 1, sy) * ixy + mixed<RO>(M(1, v.ry, v.rz), i, sy, sz) * iyz +
            mixed<RO>(M(1, v.rz, v.rx), i, sz, 1) * izx;
    r4[i] = mixed<RO>(M(2, v.rx, v.ry), i, 1, sy) * ixy + mixed<RO>(M(2, v.ry, v.rz), i, sy, sz) * iyz +
            mixed<RO>(M(2, v.rz, v.rx), i, sz, 1) * izx;
}

// ---- a7 adjoint update u += dt (D+ d + rhs) (sponge: u = A u + B dt (...)) -----------------------------------------
template <int RO, bool SPONGE>
__global__ __launch_bounds__(64) void k_adj_update(float* __restrict__ u, float* __restrict__ w,
                                                   const float* __restrict__ d1x, const float* __restrict__ d2x,
                                                   const float* __restrict__ d1y, const float* __restrict__ d2y,
                                                   const float* __restrict__ d1z, const float* __restrict__ d2z,
                                                   const float* __restrict__ r3, const float* __restrict__ r4,
                                                   const ab2* __restrict__ abx, const ab2* __restrict__ aby,
                                                   const ab2* __restrict__ abz, Geo g, float dt, float ix_, float iy_,
                                                   float iz_)
{
    int ix, iy, iz;
    long i;
    if (!point(g, ix, iy, iz, i))
        return;
    const long sy = g.ldx, sz = g.ldx * g.ldy;
    auto G = [](const float* p) { return [p](long j) { return p[j]; }; };
    const float q3 = dplus<RO>(G(d1x), i, 1) * ix_ + dplus<RO>(G(d1y), i, sy) * iy_ + dplus<RO>(G(d1z), i, sz) * iz_ +
                     r3[i];
    const float q4 = dplus<RO>(G(d2x), i, 1) * ix_ + dplus<RO>(G(d2y), i, sy) * iy_ + dplus<RO>(G(d2z), i, sz) * iz_ +
                     r4[i];
    if constexpr (SPONGE) {
        const float a = fminf(fminf(abx[ix].a, aby[iy].a), abz[iz].a);
        const float bb = fminf(fminf(abx[ix].b, aby[iy].b), abz[iz].b) * dt;
        u[i] = a * u[i] + bb * q3;
        w[i] = a * w[i] + bb * q4;
    } else {
        u[i] += q3 * dt;
        w[i] += q4 * dt;
    }
}

// ---- a10 density gradient g += -b sum_{f,a} [S D+ Dfwd + (D- S) Dfwd] ----------------------------------------------
template <int RO>
__global__ __launch_bounds__(64) void k_rho_grad(float* __restrict__ grad, const float* __restrict__ e1x,
                                                 const float* __restrict__ e2x, const float* __restrict__ e1y,
                                                 const float* __restrict__ e2y, const float* __restrict__ e1z,
                                                 const float* __restrict__ e2z, Vols v, Geo g, long gldx, long gldy,
                                                 float hx, float hy, float hz)
{
    int ix, iy, iz;
    long i;
    if (!point(g, ix, iy, iz, i))
        return;
    const long sy = g.ldx, sz = g.ldx * g.ldy;
    auto G = [](const float* p) { return [p](long j) { return p[j]; }; };
    auto S = [&](int f, const float* r) {
        return [&v, f, r](long j) {
            const Lh t = v.lh(j);
            const float ra = r[j];
            return v.rho(j) * (f == 1 ? t.lap1 + ra * ra * t.hz1 : t.lap2 + ra * ra * t.hz2);
        };
    };
    const float* e1[3] = { e1x, e1y, e1z };
    const float* e2[3] = { e2x, e2y, e2z };
    const float* r[3] = { v.rx, v.ry, v.rz };
    const long st[3] = { 1, sy, sz };
    const float h[3] = { hx, hy, hz };
    float rhs = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const auto s1 = S(1, r[a]), s2 = S(2, r[a]);
        rhs += s1(i) * (dplus<RO>(G(e1[a]), i, st[a]) * h[a]) + s2(i) * (dplus<RO>(G(e2[a]), i, st[a]) * h[a]);
        rhs += (dminus<RO>(s1, i, st[a]) * h[a]) * e1[a][i] + (dminus<RO>(s2, i, st[a]) * h[a]) * e2[a][i];
    }
    grad[((long)(iz - g.izbeg) * gldy + (iy - g.iybeg)) * gldx + (ix - g.ixbeg)] += -v.b[i] * rhs;
}

// ---- a11 / a12 constant-density (type1) forward and adjoint --------------------------------------------------------
template <int RO>
__global__ __launch_bounds__(64) void k_t1_fwd(float* __restrict__ u, float* __restrict__ w, Vols v,
                                               const ab2* __restrict__ abx, const ab2* __restrict__ aby,
                                               const ab2* __restrict__ abz, Geo g, long oldx, long oldy, float dt,
                                               float ixx, float iyy, float izz, float ixy, float iyz, float izx,
                                               int mode)
{
    int ix, iy, iz;
    long i;
    if (!point(g, ix, iy, iz, i))
        return;
    const long sy = g.ldx, sz = g.ldx * g.ldy;
    auto G = [](const float* p) { return [p](long j) { return p[j]; }; };
    auto F1 = G(v.f1);
    auto F2 = G(v.f2);
    const float a1[6] = { central2<RO>(F1, i, 1) * ixx, central2<RO>(F1, i, sy) * iyy, central2<RO>(F1, i, sz) * izz,
                          mixed<RO>(F1, i, 1, sy) * ixy, mixed<RO>(F1, i, sy, sz) * iyz, mixed<RO>(F1, i, sz, 1) * izx };
    const float a2[6] = { central2<RO>(F2, i, 1) * ixx, central2<RO>(F2, i, sy) * iyy, central2<RO>(F2, i, sz) * izz,
                          mixed<RO>(F2, i, 1, sy) * ixy, mixed<RO>(F2, i, sy, sz) * iyz, mixed<RO>(F2, i, sz, 1) * izx };
    float r3, r4;
    combine(v.m1[i], v.m2[i], v.m3[i], v.m4[i], v.rx[i], v.ry[i], v.rz[i], a1, a2, r3, r4);
    if (mode == 2) { // calc_snap: the right-hand side into a no-halo volume
        const long o = ((long)(iz - g.izbeg) * oldy + (iy - g.iybeg)) * oldx + (ix - g.ixbeg);
        u[o] = r3;
        w[o] = r4;
    } else if (mode == 1) {
        const float a = fminf(abx[ix].a, fminf(aby[iy].a, abz[iz].a));
        const float bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[iz].b));
        u[i] = a * u[i] + bb * r3 * dt;
        w[i] = a * w[i] + bb * r4 * dt;
    } else {
        u[i] += r3 * dt;
        w[i] += r4 * dt;
    }
}

template <int RO>
__global__ __launch_bounds__(64) void k_t1_adj(float* __restrict__ u, float* __restrict__ w, Vols v,
                                               const ab2* __restrict__ abx, const ab2* __restrict__ aby,
                                               const ab2* __restrict__ abz, Geo g, float dt, float ixx, float iyy,
                                               float izz, float ixy, float iyz, float izx, int use_sponge)
{
    int ix, iy, iz;
    long i;
    if (!point(g, ix, iy, iz, i))
        return;
    const long sy = g.ldx, sz = g.ldx * g.ldy;
    // T_{f,a} = lap_f + r_a^2 hz_f, T_{f,ab} = 2 r_a r_b hz_f (no rho), formed at every stencil point
    auto T = [&](int f, const float* r) {
        return [&v, f, r](long j) {
            const Lh t = v.lh(j);
            const float ra = r[j];
            return f == 1 ? t.lap1 + ra * ra * t.hz1 : t.lap2 + ra * ra * t.hz2;
        };
    };
    auto M = [&](int f, const float* ra, const float* rb) {
        return [&v, f, ra, rb](long j) {
            const Lh t = v.lh(j);
            return 2.0f * ra[j] * rb[j] * (f == 1 ? t.hz1 : t.hz2);
        };
    };
    const float q3 = central2<RO>(T(1, v.rx), i, 1) * ixx + central2<RO>(T(1, v.ry), i, sy) * iyy +
                     central2<RO>(T(1, v.rz), i, sz) * izz + mixed<RO>(M(1, v.rx, v.ry), i, 1, sy) * ixy +
                     mixed<RO>(M(1, v.ry, v.rz), i, sy, sz) * iyz + mixed<RO>(M(1, v.rz, v.rx), i, sz, 1) * izx;
    const float q4 = central2<RO>(T(2, v.rx), i, 1) * ixx + central2<RO>(T(2, v.ry), i, sy) * iyy +
                     central2<RO>(T(2, v.rz), i, sz) * izz + mixed<RO>(M(2, v.rx, v.ry), i, 1, sy) * ixy +
                     mixed<RO>(M(2, v.ry, v.rz), i, sy, sz) * iyz + mixed<RO>(M(2, v.rz, v.rx), i, sz, 1) * izx;
    if (use_sponge) {
        const float a = fminf(abx[ix].a, fminf(aby[iy].a, abz[iz].a));
        const float bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[iz].b));
        u[i] = a * u[i] + bb * q3 * dt;
        w[i] = a * w[i] + bb * q4 * dt;
    } else {
        u[i] += q3 * dt;
        w[i] += q4 * dt;
    }
}

inline dim3
grid_of(const Geo& g)
{
    return dim3((unsigned)((g.ixend - g.ixbeg + 64) / 64), (unsigned)(g.iyend - g.iybeg + 1),
                (unsigned)(g.izend - g.izbeg + 1));
}

// ghost zones of at least RO points around the interior in x and y; z is the caller's contract, as for order 8
inline bool
geo_ok(const Geo& g, int ro)
{
    return g.ixbeg >= ro && g.iybeg >= ro && g.izbeg >= ro && g.ixend >= g.ixbeg && g.iyend >= g.iybeg &&
           g.izend >= g.izbeg && g.ixend + ro < g.ldx && g.iyend + ro < g.ldy && g.iyend - g.iybeg < 65535 &&
           g.izend - g.izbeg < 65535;
}

template <class K>
int
dispatch(int ro, K&& k)
{
    switch (ro) {
        case 4: k(std::integral_constant<int, 4>{}); break;
        case 6: k(std::integral_constant<int, 6>{}); break;
        case 8: k(std::integral_constant<int, 8>{}); break;
        default: return TTI_EORDER;
    }
    return launch_status();
}

} // namespace order
} // namespace tti

using namespace tti;
using namespace tti::order;

namespace tti {

int
order_fwd_drv(int order, const float* f1, const float* f2, const float* b, float* d1x, float* d1y, float* d1z,
              float* d2x, float* d2y, float* d2z, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
              float invdx, float invdy, float invdz, int ldimx, int ldimy, void* stream)
{
    const Geo g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy };
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!geo_ok(g, order / 2))
        return TTI_EINVAL;
    return dispatch(order / 2, [&](auto ro) {
        k_fwd_drv<decltype(ro)::value><<<grid_of(g), dim3(64), 0, (hipStream_t)stream>>>(
            f1, f2, b, d1x, d1y, d1z, d2x, d2y, d2z, g, invdx, invdy, invdz);
    });
}

int
order_fwd_update(int order, float* u, float* w, const float* d1x, const float* d2x, const float* d1y,
                 const float* d2y, const float* d1z, const float* d2z, const float* f1, const float* f2,
                 const float* irho, const float* m1, const float* m2, const float* m3, const float* m4,
                 const float* rx, const float* ry, const float* rz, int ixbeg, int ixend, int iybeg, int iyend,
                 int izbeg, int izend, int ldimx, int ldimy, float dt, double dx, double dy, double dz, void* stream)
{
    const Geo g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy };
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!geo_ok(g, order / 2))
        return TTI_EINVAL;
    const Vols v{ f1, f2, m1, m2, m3, m4, rx, ry, rz, irho };
    return dispatch(order / 2, [&](auto ro) {
        k_fwd_update<decltype(ro)::value><<<grid_of(g), dim3(64), 0, (hipStream_t)stream>>>(
            u, w, d1x, d2x, d1y, d2y, d1z, d2z, v, g, dt, (float)(1.0 / dx), (float)(1.0 / dy), (float)(1.0 / dz),
            (float)(1.0 / (4 * dx * dy)), (float)(1.0 / (4 * dy * dz)), (float)(1.0 / (4 * dx * dz)));
    });
}

int
order_adj_drv(int order, float* d1x, float* d2x, float* d1y, float* d2y, float* d1z, float* d2z, float* r3,
              float* r4, const float* f1, const float* f2, const float* b, const float* m1, const float* m2,
              const float* m3, const float* m4, const float* rx, const float* ry, const float* rz, int ixbeg,
              int ixend, int iybeg, int iyend, int izbeg, int izend, float invdxy, float invdyz, float invdzx,
              int ldimx, int ldimy, float invdx, float invdy, float invdz, void* stream)
{
    const Geo g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy };
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!geo_ok(g, order / 2))
        return TTI_EINVAL;
    const Vols v{ f1, f2, m1, m2, m3, m4, rx, ry, rz, b };
    return dispatch(order / 2, [&](auto ro) {
        k_adj_drv<decltype(ro)::value><<<grid_of(g), dim3(64), 0, (hipStream_t)stream>>>(
            d1x, d2x, d1y, d2y, d1z, d2z, r3, r4, v, g, invdx, invdy, invdz, invdxy, invdyz, invdzx);
    });
}

int
order_adj_update(int order, float* u, float* w, const float* d1x, const float* d2x, const float* d1y,
                 const float* d2y, const float* d1z, const float* d2z, const float* r3, const float* r4,
                 const void* abx, const void* aby, const void* abz, int ixbeg, int ixend, int iybeg, int iyend,
                 int izbeg, int izend, int ldimx, int ldimy, float dt, double dx, double dy, double dz,
                 int sponge_active, void* stream)
{
    const Geo g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy };
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!geo_ok(g, order / 2))
        return TTI_EINVAL;
    const float ix_ = (float)(1.0 / dx), iy_ = (float)(1.0 / dy), iz_ = (float)(1.0 / dz);
    return dispatch(order / 2, [&](auto ro) {
        constexpr int RO = decltype(ro)::value;
        auto k = sponge_active ? k_adj_update<RO, true> : k_adj_update<RO, false>;
        k<<<grid_of(g), dim3(64), 0, (hipStream_t)stream>>>(u, w, d1x, d2x, d1y, d2y, d1z, d2z, r3, r4,
                                                            (const ab2*)abx, (const ab2*)aby, (const ab2*)abz, g, dt,
                                                            ix_, iy_, iz_);
    });
}

int
order_rho_grad(int order, float* grad, const float* e1x, const float* e2x, const float* e1y, const float* e2y,
               const float* e1z, const float* e2z, const float* f1, const float* f2, const float* irho,
               const float* m1, const float* m2, const float* m3, const float* m4, const float* rx, const float* ry,
               const float* rz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy,
               int grad_ldimx, int grad_ldimy, double dx, double dy, double dz, void* stream)
{
    const Geo g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy };
    if (order != 8 && order != 12 && order != 16)
        return TTI_EORDER;
    if (!geo_ok(g, order / 2))
        return TTI_EINVAL;
    const Vols v{ f1, f2, m1, m2, m3, m4, rx, ry, rz, irho };
    return dispatch(order / 2, [&](auto ro) {
        k_rho_grad<decltype(ro)::value><<<grid_of(g), dim3(64), 0, (hipStream_t)stream>>>(
            grad, e1x, e2x, e1y, e2y, e1z, e2z, v, g, (long)grad_ldimx, (long)grad_ldimy, (float)(1.0 / dx),
            (float)(1.0 / dy), (float)(1.0 / dz));
    });
}

int
order_t1_fwd(int order, float* u, float* w, const float* f1, const float* f2, const float* m1, const float* m2,
             const float* m3, const float* m4, const float* m5, const float* m6, const float* m7, const void* abx,
             const void* aby, const void* abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
             int ldimx_in, int ldimy_in, int ldimx_out, int ldimy_out, float dt, double dx, double dy, double dz,
             int mode, void* stream)
{
    const Geo g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx_in, (long)ldimy_in };
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!geo_ok(g, order / 2))
        return TTI_EINVAL;
    const Vols v{ f1, f2, m1, m2, m3, m4, m5, m6, m7, nullptr };
    return dispatch(order / 2, [&](auto ro) {
        k_t1_fwd<decltype(ro)::value><<<grid_of(g), dim3(64), 0, (hipStream_t)stream>>>(
            u, w, v, (const ab2*)abx, (const ab2*)aby, (const ab2*)abz, g, (long)ldimx_out, (long)ldimy_out, dt,
            (float)(1.0 / (dx * dx)), (float)(1.0 / (dy * dy)), (float)(1.0 / (dz * dz)), (float)(1.0 / (4 * dx * dy)),
            (float)(1.0 / (4 * dy * dz)), (float)(1.0 / (4 * dx * dz)), mode);
    });
}

int
order_t1_adj(int order, float* u, float* w, const float* f1, const float* f2, const float* m1, const float* m2,
             const float* m3, const float* m4, const float* m5, const float* m6, const float* m7, const void* abx,
             const void* aby, const void* abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
             int ldimx, int ldimy, float dt, double dx, double dy, double dz, int use_sponge, void* stream)
{
    const Geo g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy };
    if (order != 8 && order != 16)
        return TTI_EORDER;
    if (!geo_ok(g, order / 2))
        return TTI_EINVAL;
    const Vols v{ f1, f2, m1, m2, m3, m4, m5, m6, m7, nullptr };
    return dispatch(order / 2, [&](auto ro) {
        k_t1_adj<decltype(ro)::value><<<grid_of(g), dim3(64), 0, (hipStream_t)stream>>>(
            u, w, v, (const ab2*)abx, (const ab2*)aby, (const ab2*)abz, g, dt, (float)(1.0 / (dx * dx)),
            (float)(1.0 / (dy * dy)), (float)(1.0 / (dz * dz)), (float)(1.0 / (4 * dx * dy)),
            (float)(1.0 / (4 * dy * dz)), (float)(1.0 / (4 * dx * dz)), use_sponge);
    });
}

} // namespace tti
