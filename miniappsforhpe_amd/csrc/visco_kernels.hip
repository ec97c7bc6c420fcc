// Viscoacoustic memory-variable passes (kernels/kernel_base.cu:189-504): nmech relaxation mechanisms per point, the
// reference fixes nmech = 3 (kernels/kernel_base.h:27). M8[m] / M9[m] are the memory variables of sigma_H / sigma_V,
// M10[m] = alpha_m (a model volume), wsi[m] = omega_m (a constant per mechanism).
//
//   loop1 (forward, in place of loop2): sigma <- A sigma + B dt (u - sum_m omega_m M_m)
//   loop3 (forward, after loop1):       M_m  <- M_m + dt (alpha_m u - omega_m M_m)
//   loop3 with the Born fields:         the same for M_m and for the Born memory variables M1_m of u1 (alpha read once)
//   adj_loop3 (adjoint, first):         M_m  <- M_m + dt omega_m (-u - M_m)
//   adj_loop1 (adjoint, in place of loop2): sigma <- A sigma + B dt (u + sum_m alpha_m M_m)
//
// Pointwise streaming kernels over the interior: one thread per point, x-fastest rows of 256 threads per block (as
// loop2), the per-mechanism pointers and omegas passed by value. Every sum of products is an explicit fmaf chain in
// the reference's order, so the rounding does not depend on what the compiler chooses to contract.
#include "common.h"
#include "tti_hip.h"

namespace tti {
namespace visco {

constexpr int NMAX = 3;

struct Mech
{
    float* m8[NMAX];
    float* m9[NMAX];
    const float* m10[NMAX];
    float w[NMAX];
};

struct Point
{
    long i;
    float a, b;
};

__device__ __forceinline__ bool
locate(int ixbeg, int ixend, int iybeg, int izbeg, long ldx, long ldy, const ab2* __restrict__ abx,
       const ab2* __restrict__ aby, const ab2* __restrict__ abz, Point& p)
{
    const int ix = ixbeg + blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = iybeg + blockIdx.y;
    const int iz = izbeg + blockIdx.z;
    if (ix > ixend)
        return false;
    p.i = ((long)iz * ldy + iy) * ldx + ix;
    p.a = 1.0f;
    p.b = 1.0f;
    if (abx != nullptr) {
        p.a = fminf(abx[ix].a, fminf(aby[iy].a, abz[iz].a));
        p.b = fminf(abx[ix].b, fminf(aby[iy].b, abz[iz].b));
    }
    return true;
}

// loop1 (FWD = true) and adj_loop1 (FWD = false): sigma <- A sigma + B dt (u -/+ sum_m c_m M_m), c = omega (forward)
// or alpha (adjoint)
template <int NM, bool FWD>
__global__ __launch_bounds__(256) void k_loop1(float* __restrict__ f1, float* __restrict__ f2,
                                               const float* __restrict__ g1, const float* __restrict__ g2, Mech mc,
                                               const ab2* __restrict__ abx, const ab2* __restrict__ aby,
                                               const ab2* __restrict__ abz, int ixbeg, int ixend, int iybeg,
                                               int izbeg, long ldx, long ldy, float dt)
{
    Point p;
    if (!locate(ixbeg, ixend, iybeg, izbeg, ldx, ldy, abx, aby, abz, p))
        return;
    const long i = p.i;
    float rx = 0.0f, rz = 0.0f;
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const float c = FWD ? mc.w[m] : mc.m10[m][i];
        rx = fmaf(c, mc.m8[m][i], rx);
        rz = fmaf(c, mc.m9[m][i], rz);
    }
    const float bdt = p.b * dt;
    const float r1 = fmaf(bdt, FWD ? g1[i] - rx : g1[i] + rx, p.a * f1[i]);
    const float r2 = fmaf(bdt, FWD ? g2[i] - rz : g2[i] + rz, p.a * f2[i]);
    f1[i] = r1;
    f2[i] = r2;
}

// loop3: M_m <- M_m + dt (alpha_m u - omega_m M_m)
template <int NM>
__global__ __launch_bounds__(256) void k_loop3(const float* __restrict__ g1, const float* __restrict__ g2, Mech mc,
                                               int ixbeg, int ixend, int iybeg, int izbeg, long ldx, long ldy,
                                               float dt)
{
    Point p;
    if (!locate(ixbeg, ixend, iybeg, izbeg, ldx, ldy, nullptr, nullptr, nullptr, p))
        return;
    const long i = p.i;
    const float u = g1[i], w = g2[i];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const float al = mc.m10[m][i], om = mc.w[m];
        const float mx = mc.m8[m][i], mz = mc.m9[m][i];
        mc.m8[m][i] = fmaf(dt, fmaf(al, u, -(om * mx)), mx);
        mc.m9[m][i] = fmaf(dt, fmaf(al, w, -(om * mz)), mz);
    }
}

// loop3 of the reference's update_born branch (kernels/kernel_base.cu:277-330): both sets of memory variables in one
// pass, alpha_m loaded once per point; each set follows the loop3 rule, so either output equals loop3 on that set alone
template <int NM>
__global__ __launch_bounds__(256) void k_loop3_born(const float* __restrict__ g1, const float* __restrict__ g2,
                                                    const float* __restrict__ h1, const float* __restrict__ h2, Mech mc,
                                                    Mech mb, int ixbeg, int ixend, int iybeg, int izbeg, long ldx,
                                                    long ldy, float dt)
{
    Point p;
    if (!locate(ixbeg, ixend, iybeg, izbeg, ldx, ldy, nullptr, nullptr, nullptr, p))
        return;
    const long i = p.i;
    const float u = g1[i], w = g2[i], u1 = h1[i], w1 = h2[i];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const float al = mc.m10[m][i], om = mc.w[m];
        const float mx = mc.m8[m][i], mz = mc.m9[m][i];
        const float mx1 = mb.m8[m][i], mz1 = mb.m9[m][i];
        mc.m8[m][i] = fmaf(dt, fmaf(al, u, -(om * mx)), mx);
        mc.m9[m][i] = fmaf(dt, fmaf(al, w, -(om * mz)), mz);
        mb.m8[m][i] = fmaf(dt, fmaf(al, u1, -(om * mx1)), mx1);
        mb.m9[m][i] = fmaf(dt, fmaf(al, w1, -(om * mz1)), mz1);
    }
}

// adj_loop3: M_m <- M_m + dt omega_m (-u - M_m)
template <int NM>
__global__ __launch_bounds__(256) void k_adj_loop3(const float* __restrict__ g1, const float* __restrict__ g2, Mech mc,
                                                   int ixbeg, int ixend, int iybeg, int izbeg, long ldx, long ldy,
                                                   float dt)
{
    Point p;
    if (!locate(ixbeg, ixend, iybeg, izbeg, ldx, ldy, nullptr, nullptr, nullptr, p))
        return;
    const long i = p.i;
    const float u = g1[i], w = g2[i];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        const float dom = dt * mc.w[m];
        const float mx = mc.m8[m][i], mz = mc.m9[m][i];
        mc.m8[m][i] = fmaf(dom, -u - mx, mx);
        mc.m9[m][i] = fmaf(dom, -w - mz, mz);
    }
}

} // namespace visco
} // namespace tti

using namespace tti;
using namespace tti::visco;

namespace {

bool
interior_ok(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy)
{
    return ixend >= ixbeg && iyend >= iybeg && izend >= izbeg && ixbeg >= 0 && iybeg >= 0 && izbeg >= 0 &&
           ixend < ldimx && iyend < ldimy && iyend - iybeg < 65536 && izend - izbeg < 65536;
}

bool
sponge_ok(const void* a, const void* b, const void* c)
{
    return (a && b && c) || (!a && !b && !c);
}

// fills the mechanism table; false if a pointer the pass needs is null or nmech is out of range
bool
mech(int nmech, float* const* m8, float* const* m9, const float* const* m10, const float* wsi, Mech& mc)
{
    if (nmech < 1 || nmech > NMAX || !m8 || !m9)
        return false;
    for (int m = 0; m < NMAX; ++m) {
        mc.m8[m] = m < nmech ? m8[m] : nullptr;
        mc.m9[m] = m < nmech ? m9[m] : nullptr;
        mc.m10[m] = m < nmech && m10 ? m10[m] : nullptr;
        mc.w[m] = m < nmech && wsi ? wsi[m] : 0.0f;
        if (m < nmech && (!mc.m8[m] || !mc.m9[m] || (m10 && !mc.m10[m])))
            return false;
    }
    return true;
}

dim3
grid_of(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend)
{
    return dim3((ixend - ixbeg + 1 + 255) / 256, iyend - iybeg + 1, izend - izbeg + 1);
}

} // namespace

extern "C" {

int
kernel_loop1_inner(float* field1, float* field2, const float* field3, const float* field4, const float* const* Model8,
                   const float* const* Model9, const float* wsi, int nmech, const void* ABx, const void* ABy,
                   const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx,
                   int ldimy, float dt, void* stream)
{
    Mech mc;
    if (!field1 || !field2 || !field3 || !field4 || !wsi || !sponge_ok(ABx, ABy, ABz) ||
        !interior_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !mech(nmech, const_cast<float* const*>(Model8), const_cast<float* const*>(Model9), nullptr, wsi, mc))
        return TTI_EINVAL;
    const dim3 g = grid_of(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    auto go = [&](auto kernel) {
        kernel<<<g, dim3(256), 0, (hipStream_t)stream>>>(field1, field2, field3, field4, mc, (const ab2*)ABx,
                                                         (const ab2*)ABy, (const ab2*)ABz, ixbeg, ixend, iybeg, izbeg,
                                                         (long)ldimx, (long)ldimy, dt);
    };
    nmech == 1 ? go(k_loop1<1, true>) : nmech == 2 ? go(k_loop1<2, true>) : go(k_loop1<3, true>);
    return launch_status();
}

int
kernel_loop3_inner(const float* field3, const float* field4, float* const* Model8, float* const* Model9,
                   const float* const* Model10, const float* wsi, int nmech, int ixbeg, int ixend, int iybeg,
                   int iyend, int izbeg, int izend, int ldimx, int ldimy, float dt, void* stream)
{
    Mech mc;
    if (!field3 || !field4 || !wsi || !Model10 ||
        !interior_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !mech(nmech, Model8, Model9, Model10, wsi, mc))
        return TTI_EINVAL;
    const dim3 g = grid_of(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    auto go = [&](auto kernel) {
        kernel<<<g, dim3(256), 0, (hipStream_t)stream>>>(field3, field4, mc, ixbeg, ixend, iybeg, izbeg, (long)ldimx,
                                                         (long)ldimy, dt);
    };
    nmech == 1 ? go(k_loop3<1>) : nmech == 2 ? go(k_loop3<2>) : go(k_loop3<3>);
    return launch_status();
}

int
kernel_loop3_inner_born(const float* field3, const float* field4, const float* field31, const float* field41,
                        float* const* Model8, float* const* Model9, float* const* Model81, float* const* Model91,
                        const float* const* Model10, const float* wsi, int nmech, int ixbeg, int ixend, int iybeg,
                        int iyend, int izbeg, int izend, int ldimx, int ldimy, float dt, void* stream)
{
    Mech mc, mb;
    if (!field3 || !field4 || !field31 || !field41 || !wsi || !Model10 ||
        !interior_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !mech(nmech, Model8, Model9, Model10, wsi, mc) || !mech(nmech, Model81, Model91, nullptr, nullptr, mb))
        return TTI_EINVAL;
    const dim3 g = grid_of(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    auto go = [&](auto kernel) {
        kernel<<<g, dim3(256), 0, (hipStream_t)stream>>>(field3, field4, field31, field41, mc, mb, ixbeg, ixend, iybeg,
                                                         izbeg, (long)ldimx, (long)ldimy, dt);
    };
    nmech == 1 ? go(k_loop3_born<1>) : nmech == 2 ? go(k_loop3_born<2>) : go(k_loop3_born<3>);
    return launch_status();
}

int
adj_kernel_loop1_inner(float* field1, float* field2, const float* field3, const float* field4,
                       const float* const* Model8, const float* const* Model9, const float* const* Model10, int nmech,
                       const void* ABx, const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend,
                       int izbeg, int izend, int ldimx, int ldimy, float dt, void* stream)
{
    Mech mc;
    if (!field1 || !field2 || !field3 || !field4 || !Model10 || !sponge_ok(ABx, ABy, ABz) ||
        !interior_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !mech(nmech, const_cast<float* const*>(Model8), const_cast<float* const*>(Model9), Model10, nullptr, mc))
        return TTI_EINVAL;
    const dim3 g = grid_of(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    auto go = [&](auto kernel) {
        kernel<<<g, dim3(256), 0, (hipStream_t)stream>>>(field1, field2, field3, field4, mc, (const ab2*)ABx,
                                                         (const ab2*)ABy, (const ab2*)ABz, ixbeg, ixend, iybeg, izbeg,
                                                         (long)ldimx, (long)ldimy, dt);
    };
    nmech == 1 ? go(k_loop1<1, false>) : nmech == 2 ? go(k_loop1<2, false>) : go(k_loop1<3, false>);
    return launch_status();
}

int
adj_kernel_loop3_inner(const float* field3, const float* field4, float* const* Model8, float* const* Model9,
                       const float* wsi, int nmech, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                       int ldimx, int ldimy, float dt, void* stream)
{
    Mech mc;
    if (!field3 || !field4 || !wsi || !interior_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !mech(nmech, Model8, Model9, nullptr, wsi, mc))
        return TTI_EINVAL;
    const dim3 g = grid_of(ixbeg, ixend, iybeg, iyend, izbeg, izend);
    auto go = [&](auto kernel) {
        kernel<<<g, dim3(256), 0, (hipStream_t)stream>>>(field3, field4, mc, ixbeg, ixend, iybeg, izbeg, (long)ldimx,
                                                         (long)ldimy, dt);
    };
    nmech == 1 ? go(k_adj_loop3<1>) : nmech == 2 ? go(k_adj_loop3<2>) : go(k_adj_loop3<3>);
    return launch_status();
}

} // extern "C"
