// Single-volume fused TTI time steps (one launch per step), receiver/source helpers and the snapshot decoder, with
// their C ABI (include/tti_fused.h). The tile bodies are in fused_tiles.h.
#include "fused_tiles.h"
#include "tti_hip.h"
#include "type2_internal.h"

#include <initializer_list>

namespace tti {

namespace fused {

template <int FBY, int SNAP, bool SPONGE, bool SIG = false>
__global__ __launch_bounds__(BX* FBY, FBY <= 4 ? 2 : 1) void k_fused_fwd( // at least 2 waves per SIMD: <= 256 VGPRs
    const float* __restrict__ s1, const float* __restrict__ s2, float* __restrict__ s1o, float* __restrict__ s2o,
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ m1, const float* __restrict__ m2,
    const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ vrx,
    const float* __restrict__ vry, const float* __restrict__ vrz, const float* __restrict__ b,
    const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, Box g, float dt,
    float invdx, float invdy, float invdz, float i4xy, float i4yz, float i4zx, unsigned src_off, float src_val,
    Snap snap, int bz, int ntx, int nty, int ntz, int strip)
{
    int bxi, byi, bzi;
    if (!decode_tile(ntx, nty, ntz, strip, bxi, byi, bzi))
        return;
    fwd_tile<FBY, SNAP, SPONGE, SIG>(s1, s2, s1o, s2o, u, w, m1, m2, m3, m4, vrx, vry, vrz, b, abx, aby, abz, g, dt, invdx,
                                invdy, invdz, i4xy, i4yz, i4zx, src_off, src_val, snap, bz, bxi, byi, bzi);
}

template <int SNAP, bool SPONGE>
__global__ __launch_bounds__(ADJ_AX* ADJ_AY, ADJ_AX* ADJ_AY <= 256 ? 2 : 1) void k_fused_adj(
    const float* __restrict__ a1, const float* __restrict__ a2, float* __restrict__ a1o, float* __restrict__ a2o,
    float* __restrict__ u, float* __restrict__ w, const float* __restrict__ m1, const float* __restrict__ m2,
    const float* __restrict__ m3, const float* __restrict__ m4, const float* __restrict__ vrx,
    const float* __restrict__ vry, const float* __restrict__ vrz, const float* __restrict__ b,
    const ab2* __restrict__ abx, const ab2* __restrict__ aby, const ab2* __restrict__ abz, Box g, float dt,
    float invdx, float invdy, float invdz, float i4xy, float i4yz, float i4zx, Snap snap, double* __restrict__ grad,
    int bz, int ntx, int nty, int ntz, int strip)
{
    int bxi, byi, bzi;
    if (!decode_tile(ntx, nty, ntz, strip, bxi, byi, bzi))
        return;
    adj_tile<SNAP, SPONGE>(a1, a2, a1o, a2o, u, w, m1, m2, m3, m4, vrx, vry, vrz, b, abx, aby, abz, g, dt, invdx, invdy,
                           invdz, i4xy, i4yz, i4zx, snap, grad, bz, bxi, byi, bzi);
}

// ------------------------------------------------------------------------------------------------------------------
// Receivers and sources
// ------------------------------------------------------------------------------------------------------------------

// out[k] = -(2 s1 + s2)/3 at idx[k] (pressure recorded at receivers, simulator2/simulator2SetUp.cu:31-42)
__global__ __launch_bounds__(256) void k_gather_pressure(const float* __restrict__ s1, const float* __restrict__ s2,
                                                         const long* __restrict__ idx, int n, float* __restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        out[k] = pressure_at(s1, s2, idx[k]);
}

// Transpose of the pressure recording, applied after an adjoint step (inject_point). Receivers must be distinct
// points.
__global__ __launch_bounds__(256) void k_inject_residual(float* __restrict__ u, float* __restrict__ w,
                                                         float* __restrict__ a1o, float* __restrict__ a2o,
                                                         const long* __restrict__ idx, const int* __restrict__ ijk,
                                                         int n, const float* __restrict__ res, float dt,
                                                         const ab2* __restrict__ abx, const ab2* __restrict__ aby,
                                                         const ab2* __restrict__ abz, Box g, Snap snap,
                                                         double* __restrict__ grad)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        inject_point(u, w, a1o, a2o, idx, ijk, k, res, dt, abx, aby, abz, g, snap, grad);
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
    out[k] = snap_load(mode, arr, scales, row * sldx + x, row * (sldx / SNAP_SEG) + x / SNAP_SEG);
}

// Decode a snapshot field into the interior of a padded FP32 volume (ghosts untouched).
__global__ __launch_bounds__(256) void k_snap_decode_vol(int mode, const void* __restrict__ arr,
                                                         const float* __restrict__ scales, long sldx, long sldy,
                                                         int nx, int ny, int nz, float* __restrict__ out, long ldx,
                                                         long ldy, int ixbeg, int iybeg, int izbeg)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= nx)
        return;
    const long row = (long)z * sldy + y;
    out[((long)(izbeg + z) * ldy + (iybeg + y)) * ldx + (ixbeg + x)] =
        snap_load(mode, arr, scales, row * sldx + x, row * (sldx / SNAP_SEG) + x / SNAP_SEG);
}

} // namespace fused
} // namespace tti

// ======================================================================================================================
// C ABI (include/tti_fused.h)
// ======================================================================================================================
using namespace tti;
using namespace tti::fused;

namespace {

struct Launch
{
    dim3 grid;
    int ntx, nty, ntz, bz;
};

Launch
plan(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int bz, int bx, int by)
{
    Launch L;
    const int nx = ixend - ixbeg + 1, ny = iyend - iybeg + 1, nz = izend - izbeg + 1;
    L.bz = bz > 0 ? bz : 128;
    L.ntx = (nx + bx - 1) / bx;
    L.nty = (ny + by - 1) / by;
    L.ntz = (nz + L.bz - 1) / L.bz;
    const int nt = L.ntx * L.nty * L.ntz;
    L.grid = dim3(((nt + 7) / 8) * 8); // multiple of 8 so the XCD round-robin decode covers every tile
    return L;
}

// Status of the step's launch; with a free surface, then the stress-free mirror of the new sigma (kernel_loop4_inner:
// plane izbeg zeroed, planes above it the negated mirror images), as the reference applies after each time update
int
finish_step(float* s1, float* s2, int ldimx, int ldimy, int izbeg, int free_surface, void* stream)
{
    const int rc = launch_status();
    if (rc != TTI_OK || !free_surface)
        return rc;
    return kernel_loop4_inner(s1, s2, ldimx, ldimy, izbeg, stream);
}

} // namespace

extern "C" {

int
tti_fused_fwd_step(const float* field1, const float* field2, float* field1_out, float* field2_out, float* field3,
                   float* field4, const float* model1, const float* model2, const float* model3, const float* model4,
                   const float* rx, const float* ry, const float* rz, const float* Model11, const void* ABx,
                   const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                   int ldimx, int ldimy, float dt, double dx, double dy, double dz, long src_index, float src_value,
                   const tti_snapshot* snap, int free_surface, int zblock, int strip, void* stream)
{
    if (!field1 || !field2 || !field1_out || !field2_out || !field3 || !field4 || !model1 || !model2 || !model3 ||
        !model4 || !rx || !ry || !rz || !Model11 || field1 == field1_out || field2 == field2_out)
        return TTI_EINVAL;
    if (!box_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !snap_ok(snap, ixend - ixbeg + 1, iyend - iybeg + 1, izend - izbeg + 1))
        return TTI_EINVAL;
    if ((ABx || ABy || ABz) && !(ABx && ABy && ABz))
        return TTI_EINVAL;
    if (!aligned16({ field1, field2, Model11 }))
        return TTI_EINVAL;
    if (!small_volume(ldimx, ldimy, izend))
        return TTI_EINVAL;
    if (src_index >= (long)ldimx * ldimy * (izend + R + 1))
        return TTI_EINVAL;
    const unsigned src_off = src_index < 0 ? 0xffffffffu : (unsigned)(src_index * 4);
    const Box g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy, free_surface != 0 };
    const Launch L = plan(ixbeg, ixend, iybeg, iyend, izbeg, izend, zblock, BX, FWD_BY);
    const Snap sn = to_snap(snap);
    auto go = [&](auto kernel) {
        kernel<<<L.grid, dim3(BX, FWD_BY), 0, (hipStream_t)stream>>>(
            field1, field2, field1_out, field2_out, field3, field4, model1, model2, model3, model4, rx, ry, rz,
            Model11, (const ab2*)ABx, (const ab2*)ABy, (const ab2*)ABz, g, dt, (float)(1.0 / dx), (float)(1.0 / dy),
            (float)(1.0 / dz), (float)(1.0 / (4 * dx * dy)), (float)(1.0 / (4 * dy * dz)),
            (float)(1.0 / (4 * dx * dz)), src_off, src_value, sn, L.bz, L.ntx, L.nty, L.ntz, strip);
    };
    const bool sp = ABx != nullptr;
    constexpr int F = FWD_BY;
    if (sn.c != nullptr) { // sigma of the step stored as well (density image)
        switch (sn.mode) {
            case TTI_SNAP_FP32: sp ? go(k_fused_fwd<F, TTI_SNAP_FP32, true, true>) : go(k_fused_fwd<F, TTI_SNAP_FP32, false, true>); break;
            case TTI_SNAP_FP16: sp ? go(k_fused_fwd<F, TTI_SNAP_FP16, true, true>) : go(k_fused_fwd<F, TTI_SNAP_FP16, false, true>); break;
            case TTI_SNAP_BF16: sp ? go(k_fused_fwd<F, TTI_SNAP_BF16, true, true>) : go(k_fused_fwd<F, TTI_SNAP_BF16, false, true>); break;
            default: return TTI_EINVAL;
        }
        return finish_step(field1_out, field2_out, ldimx, ldimy, izbeg, free_surface, stream);
    }
    switch (sn.mode) {
        case TTI_SNAP_FP32: sp ? go(k_fused_fwd<F, TTI_SNAP_FP32, true>) : go(k_fused_fwd<F, TTI_SNAP_FP32, false>); break;
        case TTI_SNAP_FP16: sp ? go(k_fused_fwd<F, TTI_SNAP_FP16, true>) : go(k_fused_fwd<F, TTI_SNAP_FP16, false>); break;
        case TTI_SNAP_BF16: sp ? go(k_fused_fwd<F, TTI_SNAP_BF16, true>) : go(k_fused_fwd<F, TTI_SNAP_BF16, false>); break;
        default: sp ? go(k_fused_fwd<F, TTI_SNAP_NONE, true>) : go(k_fused_fwd<F, TTI_SNAP_NONE, false>); break;
    }
    return finish_step(field1_out, field2_out, ldimx, ldimy, izbeg, free_surface, stream);
}

int
tti_fused_adj_step(const float* field1, const float* field2, float* field1_out, float* field2_out, float* field3,
                   float* field4, const float* model1, const float* model2, const float* model3, const float* model4,
                   const float* rx, const float* ry, const float* rz, const float* Model11, const void* ABx,
                   const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                   int ldimx, int ldimy, float dt, double dx, double dy, double dz, const tti_snapshot* snap,
                   double* grad, int free_surface, int zblock, int strip, void* stream)
{
    if (!field1 || !field2 || !field1_out || !field2_out || !field3 || !field4 || !model1 || !model2 || !model3 ||
        !model4 || !rx || !ry || !rz || !Model11 || field1 == field1_out || field2 == field2_out)
        return TTI_EINVAL;
    if (!box_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) ||
        !snap_ok(snap, ixend - ixbeg + 1, iyend - iybeg + 1, izend - izbeg + 1))
        return TTI_EINVAL;
    if ((ABx || ABy || ABz) && !(ABx && ABy && ABz))
        return TTI_EINVAL;
    if (!aligned16({ field1, field2, model1, model2, model3, model4, rx, ry, rz, Model11 }))
        return TTI_EINVAL;
    if (!small_volume(ldimx, ldimy, izend))
        return TTI_EINVAL;
    const Box g{ ixbeg, ixend, iybeg, iyend, izbeg, izend, (long)ldimx, (long)ldimy, free_surface != 0 };
    const Launch L = plan(ixbeg, ixend, iybeg, iyend, izbeg, izend, zblock, ADJ_AX, ADJ_AY);
    const Snap sn = to_snap(snap);
    auto go = [&](auto kernel) {
        kernel<<<L.grid, dim3(ADJ_AX, ADJ_AY), 0, (hipStream_t)stream>>>(
            field1, field2, field1_out, field2_out, field3, field4, model1, model2, model3, model4, rx, ry, rz,
            Model11, (const ab2*)ABx, (const ab2*)ABy, (const ab2*)ABz, g, dt, (float)(1.0 / dx), (float)(1.0 / dy),
            (float)(1.0 / dz), (float)(1.0 / (4 * dx * dy)), (float)(1.0 / (4 * dy * dz)),
            (float)(1.0 / (4 * dx * dz)), sn, grad, L.bz, L.ntx, L.nty, L.ntz, strip);
    };
    const bool sp = ABx != nullptr;
    switch (grad != nullptr ? sn.mode : TTI_SNAP_NONE) {
        case TTI_SNAP_FP32: sp ? go(k_fused_adj<TTI_SNAP_FP32, true>) : go(k_fused_adj<TTI_SNAP_FP32, false>); break;
        case TTI_SNAP_FP16: sp ? go(k_fused_adj<TTI_SNAP_FP16, true>) : go(k_fused_adj<TTI_SNAP_FP16, false>); break;
        case TTI_SNAP_BF16: sp ? go(k_fused_adj<TTI_SNAP_BF16, true>) : go(k_fused_adj<TTI_SNAP_BF16, false>); break;
        default: sp ? go(k_fused_adj<TTI_SNAP_NONE, true>) : go(k_fused_adj<TTI_SNAP_NONE, false>); break;
    }
    return finish_step(field1_out, field2_out, ldimx, ldimy, izbeg, free_surface, stream);
}

int
tti_rho_image(const float* field1, const float* field2, const float* model1, const float* model2, const float* model3,
              const float* model4, const float* rx, const float* ry, const float* rz, const float* Model11, int ixbeg,
              int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy, double dx, double dy,
              double dz, const tti_snapshot* snap, float* work, double* grad_rho, void* stream)
{
    if (!field1 || !field2 || !model1 || !model2 || !model3 || !model4 || !rx || !ry || !rz || !Model11 || !work ||
        !grad_rho || !snap || snap->mode == TTI_SNAP_NONE || !snap->field1)
        return TTI_EINVAL;
    const int nx = ixend - ixbeg + 1, ny = iyend - iybeg + 1, nz = izend - izbeg + 1;
    if (!box_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy) || !snap_ok(snap, nx, ny, nz) ||
        ny > 65535 || nz > 65535)
        return TTI_EINVAL;
    const long vol = (long)ldimx * ldimy * (izend + R + 1);
    float* s[2] = { work, work + vol };
    float* d[6];
    for (int k = 0; k < 6; ++k)
        d[k] = work + (2 + k) * vol;
    hipStream_t st = (hipStream_t)stream;
    const void* arr[2] = { snap->field1, snap->field2 };
    const float* sc[2] = { snap->scale1, snap->scale2 };
    for (int f = 0; f < 2; ++f) {
        k_snap_decode_vol<<<dim3((nx + 255) / 256, ny, nz), dim3(256), 0, st>>>(
            snap->mode, arr[f], sc[f], snap->ldimx, snap->ldimy, nx, ny, nz, s[f], ldimx, ldimy, ixbeg, iybeg, izbeg);
        const int rc = launch_status();
        if (rc != TTI_OK)
            return rc;
    }
    // snaps_loop: the forward derivatives of the snapshot sigma
    const int rc = launch_fwd_kernel2_drv(s[0], s[1], Model11, d[0], d[1], d[2], d[3], d[4], d[5], ixbeg, ixend, iybeg,
                                          iyend, izbeg, izend, (float)(1.0 / dx), (float)(1.0 / dy), (float)(1.0 / dz),
                                          ldimx, ldimy, 8, stream);
    if (rc != TTI_OK)
        return rc;
    // update_grad_loop_2 into the FP64 image (d1x d2x d1y d2y d1z d2z order of the launcher)
    return rho_gradient_f64(grad_rho, d[0], d[3], d[1], d[4], d[2], d[5], field1, field2, Model11, model1, model2,
                            model3, model4, rx, ry, rz, ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, nx, ny,
                            dx, dy, dz, stream);
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
    if (!snap_ok(snap, ixend - ixbeg + 1, iyend - iybeg + 1, izend - izbeg + 1))
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
    if (!snap || snap->mode == TTI_SNAP_NONE || !snap_ok(snap, nx, ny, nz) || !out || nx < 1 || ny < 1 || nz < 1 ||
        which < 1 || which > 4 || (which < 3 && !snap->field1))
        return TTI_EINVAL;
    const long n = (long)nx * ny * nz;
    const void* arrs[4] = { snap->field1, snap->field2, snap->field3, snap->field4 };
    const float* scs[4] = { snap->scale1, snap->scale2, snap->scale3, snap->scale4 };
    const void* arr = arrs[which - 1];
    const float* sc = scs[which - 1];
    k_snap_decode<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream>>>(
        snap->mode, arr, sc, (long)snap->ldimx, (long)snap->ldimy, nx, ny, nz, out);
    return launch_status();
}

} // extern "C"
