// Shared definitions for the gfx950 TTI kernels: FD coefficients, grid indexing, launch-status plumbing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Return codes of every C entry point (0 = success).
#define TTI_OK 0
#define TTI_EINVAL (-1)    // inconsistent sizes / ranges / null pointers
#define TTI_EORDER (-2)    // unsupported FD order
#define TTI_EHIP_BASE (-1000) // -1000 - hipError_t

namespace tti {

// 8th-order least-squares coefficients, FP32 as in the reference tables (kernels/fd_coef.h:467-470 first-order
// staggered c1..c4; kernels/fd_coef.h:517-521 second-order central e0..e4). The staggered set equals the Taylor set
// 1225/1024, -245/3072, 49/5120, -5/7168 to FP32 rounding.
constexpr float C1 = 1.19628906250000e+00f;
constexpr float C2 = -7.97526041666667e-02f;
constexpr float C3 = 9.57031250000000e-03f;
constexpr float C4 = -6.97544642857143e-04f;
constexpr float E0 = -0.2903117E+1f;
constexpr float E1 = 0.1645351E+1f;
constexpr float E2 = -0.2236606E+0f;
constexpr float E3 = 0.3265851E-1f;
constexpr float E4 = -0.2790073E-2f;

__host__ __device__ constexpr float
cstg(int n)
{
    return n == 1 ? C1 : n == 2 ? C2 : n == 3 ? C3 : n == 4 ? C4 : 0.0f;
}
__host__ __device__ constexpr float
ecen(int n)
{
    return n == 0 ? E0 : n == 1 ? E1 : n == 2 ? E2 : n == 3 ? E3 : n == 4 ? E4 : 0.0f;
}

constexpr int R = 4; // stencil radius of order 8

struct ab2
{
    float a, b;
}; // sponge coefficient pair {A, B}, same layout as a float2

inline int
hip_status(hipError_t e)
{
    return e == hipSuccess ? TTI_OK : (TTI_EHIP_BASE - (int)e);
}

// Status of the last launch on this thread.
inline int
launch_status()
{
    return hip_status(hipGetLastError());
}

} // namespace tti
