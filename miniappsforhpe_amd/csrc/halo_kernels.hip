// Grid-layer data movement on one GPU: strided box copies (volume <-> volume, volume <-> contiguous buffer), the
// halo pack/unpack of one face, the local two-way ghost exchange between adjacent sub-volumes, a periodic self
// exchange (the stand-in for a neighbour when one GPU holds the whole subdomain), and a fill kernel.
//
// Reference counterparts: copy_and_restride_data / set (infrastructure/cart_volume_gpu_kernels.cu:23-121),
// copy_2way_X/Y/Z and pack_unpack_X/Y/Z (infrastructure/haloManager3D_gpu_kernels.cuh:14-472). Face extents follow
// the reference's buffer sizes (infrastructure/haloManager3D_gpu_kernels.cuh:271-278, :368-375, :463-470): an x face
// spans the interior y-z range, a y face spans x including its ghosts and the interior z range, a z face spans x and
// y including their ghosts, so exchanging X, then Y, then Z also fills edges and corners.
#include "common.h"
#include "tti_grid.h"

#include <algorithm>
#include <mutex>

namespace tti {

struct BoxDesc
{
    long ldx, ldy; // leading dims (x row length, rows per plane)
    int x0, y0, z0;
};

// dst(box) = src(box), box of nx x ny x nz points; one thread per point, x fastest (coalesced on both sides).
__global__ __launch_bounds__(256) void k_copy_box(const float* __restrict__ src, BoxDesc s, float* __restrict__ dst,
                                                  BoxDesc d, int nx, int ny, int nz)
{
    const long n = (long)nx * ny * nz;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const int x = (int)(k % nx);
        const long r = k / nx;
        const int y = (int)(r % ny), z = (int)(r / ny);
        dst[((long)(d.z0 + z) * d.ldy + (d.y0 + y)) * d.ldx + (d.x0 + x)] =
            src[((long)(s.z0 + z) * s.ldy + (s.y0 + y)) * s.ldx + (s.x0 + x)];
    }
}

__global__ __launch_bounds__(256) void k_fill_box(float* __restrict__ v, BoxDesc d, int nx, int ny, int nz, float value)
{
    const long n = (long)nx * ny * nz;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const int x = (int)(k % nx);
        const long r = k / nx;
        const int y = (int)(r % ny), z = (int)(r / ny);
        v[((long)(d.z0 + z) * d.ldy + (d.y0 + y)) * d.ldx + (d.x0 + x)] = value;
    }
}

__global__ __launch_bounds__(256) void k_fill(float* __restrict__ v, long n, float value)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        v[k] = value;
}

// Batched box copies: blockIdx.y selects the box, the x blocks stride over its points (x fastest).
__global__ __launch_bounds__(256) void k_copy_boxes(const tti_box_copy* __restrict__ boxes)
{
    const tti_box_copy b = boxes[blockIdx.y];
    const unsigned nx = (unsigned)b.nx, ny = (unsigned)b.ny;
    const unsigned n = nx * ny * (unsigned)b.nz; // < 2^31, checked at plan creation
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const unsigned x = k % nx, r = k / nx;
        const unsigned y = r % ny, z = r / ny;
        b.dst[((long)(b.dz0 + (int)z) * b.dldy + (b.dy0 + (int)y)) * b.dldx + (b.dx0 + (int)x)] =
            b.src[((long)(b.sz0 + (int)z) * b.sldy + (b.sy0 + (int)y)) * b.sldx + (b.sx0 + (int)x)];
    }
}

} // namespace tti

using namespace tti;

namespace {

int
launch_copy(const float* src, BoxDesc s, float* dst, BoxDesc d, int nx, int ny, int nz, void* stream)
{
    const long n = (long)nx * ny * nz;
    if (n == 0)
        return TTI_OK;
    const long blocks = std::min<long>((n + 255) / 256, 65536);
    k_copy_box<<<dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream>>>(src, s, dst, d, nx, ny, nz);
    return launch_status();
}

struct Face
{
    int x0, y0, z0, nx, ny, nz; // region in array coordinates
};

// Interior slab (send = true) or ghost slab (send = false) of width r on `side` (0 = low, 1 = high) of `axis`.
bool
face(int axis, int side, bool send, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int r, Face& f)
{
    if (axis < 0 || axis > 2 || (side != 0 && side != 1) || r < 1)
        return false;
    const int lo[3] = { ixbeg, iybeg, izbeg }, hi[3] = { ixend, iyend, izend };
    int b[3], n[3];
    for (int a = 0; a < 3; ++a) {
        // x face: interior y, z; y face: x with ghosts, interior z; z face: x and y with ghosts
        const bool with_ghosts = a < axis;
        b[a] = with_ghosts ? lo[a] - r : lo[a];
        n[a] = hi[a] - lo[a] + 1 + (with_ghosts ? 2 * r : 0);
    }
    if (hi[axis] - lo[axis] + 1 < r)
        return false;
    if (send)
        b[axis] = side == 0 ? lo[axis] : hi[axis] - r + 1;
    else
        b[axis] = side == 0 ? lo[axis] - r : hi[axis] + 1;
    n[axis] = r;
    f = Face{ b[0], b[1], b[2], n[0], n[1], n[2] };
    return true;
}

bool
dims_ok(int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy, int r)
{
    return ixbeg >= r && iybeg >= r && izbeg >= r && ixend >= ixbeg && iyend >= iybeg && izend >= izbeg &&
           ixend + r < ldimx && iyend + r < ldimy;
}

} // namespace

extern "C" {

int
tti_copy_box(const float* src, int sldx, int sldy, int sx0, int sy0, int sz0, float* dst, int dldx, int dldy, int dx0,
             int dy0, int dz0, int nx, int ny, int nz, void* stream)
{
    if (!src || !dst || nx < 0 || ny < 0 || nz < 0 || sx0 < 0 || sy0 < 0 || sz0 < 0 || dx0 < 0 || dy0 < 0 ||
        dz0 < 0 || sx0 + nx > sldx || dx0 + nx > dldx || sy0 + ny > sldy || dy0 + ny > dldy)
        return TTI_EINVAL;
    return launch_copy(src, BoxDesc{ sldx, sldy, sx0, sy0, sz0 }, dst, BoxDesc{ dldx, dldy, dx0, dy0, dz0 }, nx, ny,
                       nz, stream);
}

int
tti_fill(float* v, long n, float value, void* stream)
{
    if (n < 0 || (n > 0 && !v))
        return TTI_EINVAL;
    if (n == 0)
        return TTI_OK;
    const long blocks = std::min<long>((n + 255) / 256, 65536);
    k_fill<<<dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream>>>(v, n, value);
    return launch_status();
}

int
tti_fill_box(float* v, int ldx, int ldy, int x0, int y0, int z0, int nx, int ny, int nz, float value, void* stream)
{
    if (!v || nx < 0 || ny < 0 || nz < 0 || x0 < 0 || y0 < 0 || z0 < 0 || x0 + nx > ldx || y0 + ny > ldy)
        return TTI_EINVAL;
    const long n = (long)nx * ny * nz;
    if (n == 0)
        return TTI_OK;
    const long blocks = std::min<long>((n + 255) / 256, 65536);
    k_fill_box<<<dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream>>>(v, BoxDesc{ ldx, ldy, x0, y0, z0 }, nx,
                                                                             ny, nz, value);
    return launch_status();
}

long
tti_halo_face_size(int axis, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int r)
{
    Face f;
    if (!face(axis, 0, true, ixbeg, ixend, iybeg, iyend, izbeg, izend, r, f))
        return -1;
    return (long)f.nx * f.ny * f.nz;
}

int
tti_halo_pack(const float* vol, int axis, int side, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
              int ldimx, int ldimy, int r, float* buf, void* stream)
{
    Face f;
    if (!vol || !buf || !dims_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, r) ||
        !face(axis, side, true, ixbeg, ixend, iybeg, iyend, izbeg, izend, r, f))
        return TTI_EINVAL;
    return launch_copy(vol, BoxDesc{ ldimx, ldimy, f.x0, f.y0, f.z0 }, buf, BoxDesc{ f.nx, f.ny, 0, 0, 0 }, f.nx, f.ny,
                       f.nz, stream);
}

int
tti_halo_unpack(float* vol, int axis, int side, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                int ldimx, int ldimy, int r, const float* buf, void* stream)
{
    Face f;
    if (!vol || !buf || !dims_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, r) ||
        !face(axis, side, false, ixbeg, ixend, iybeg, iyend, izbeg, izend, r, f))
        return TTI_EINVAL;
    return launch_copy(buf, BoxDesc{ f.nx, f.ny, 0, 0, 0 }, vol, BoxDesc{ ldimx, ldimy, f.x0, f.y0, f.z0 }, f.nx, f.ny,
                       f.nz, stream);
}

int
tti_halo_exchange_local(float* lo_vol, float* hi_vol, int axis, int ixbeg, int ixend, int iybeg, int iyend, int izbeg,
                        int izend, int ldimx, int ldimy, int r, void* stream)
{
    Face s_lo, g_hi, s_hi, g_lo;
    if (!lo_vol || !hi_vol || lo_vol == hi_vol || !dims_ok(ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, r) ||
        !face(axis, 1, true, ixbeg, ixend, iybeg, iyend, izbeg, izend, r, s_lo) ||
        !face(axis, 0, false, ixbeg, ixend, iybeg, iyend, izbeg, izend, r, g_hi) ||
        !face(axis, 0, true, ixbeg, ixend, iybeg, iyend, izbeg, izend, r, s_hi) ||
        !face(axis, 1, false, ixbeg, ixend, iybeg, iyend, izbeg, izend, r, g_lo))
        return TTI_EINVAL;
    // lo's high interior slab -> hi's low ghost; hi's low interior slab -> lo's high ghost (copy_2way)
    int rc = launch_copy(lo_vol, BoxDesc{ ldimx, ldimy, s_lo.x0, s_lo.y0, s_lo.z0 }, hi_vol,
                         BoxDesc{ ldimx, ldimy, g_hi.x0, g_hi.y0, g_hi.z0 }, s_lo.nx, s_lo.ny, s_lo.nz, stream);
    if (rc != TTI_OK)
        return rc;
    return launch_copy(hi_vol, BoxDesc{ ldimx, ldimy, s_hi.x0, s_hi.y0, s_hi.z0 }, lo_vol,
                       BoxDesc{ ldimx, ldimy, g_lo.x0, g_lo.y0, g_lo.z0 }, s_hi.nx, s_hi.ny, s_hi.nz, stream);
}

int
tti_halo_fill_periodic(float* vol, int axis, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                       int ldimx, int ldimy, int r, float* buf, void* stream)
{
    // buf holds two faces: [0, n) the high interior slab, [n, 2n) the low one
    const long n = tti_halo_face_size(axis, ixbeg, ixend, iybeg, iyend, izbeg, izend, r);
    if (n < 0 || !buf)
        return TTI_EINVAL;
    int rc = tti_halo_pack(vol, axis, 1, ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, r, buf, stream);
    if (rc == TTI_OK)
        rc = tti_halo_pack(vol, axis, 0, ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, r, buf + n, stream);
    if (rc == TTI_OK)
        rc = tti_halo_unpack(vol, axis, 0, ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, r, buf, stream);
    if (rc == TTI_OK)
        rc = tti_halo_unpack(vol, axis, 1, ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy, r, buf + n, stream);
    return rc;
}

} // extern "C"

struct tti_copy_plan
{
    tti_box_copy* dev = nullptr;
    int n = 0;
    long max_points = 0;
};

namespace {

bool
box_ok(const tti_box_copy& b)
{
    if (!b.src || !b.dst || b.nx < 0 || b.ny < 0 || b.nz < 0 || b.sldx <= 0 || b.sldy <= 0 || b.dldx <= 0 ||
        b.dldy <= 0 || b.sx0 < 0 || b.sy0 < 0 || b.sz0 < 0 || b.dx0 < 0 || b.dy0 < 0 || b.dz0 < 0 ||
        b.sx0 + b.nx > b.sldx || b.dx0 + b.nx > b.dldx || b.sy0 + b.ny > b.sldy || b.dy0 + b.ny > b.dldy)
        return false;
    const long n = (long)b.nx * b.ny * b.nz;
    if (n >= (1L << 31))
        return false;
    if (n == 0)
        return true;
    const long slast = ((long)(b.sz0 + b.nz - 1) * b.sldy + (b.sy0 + b.ny - 1)) * b.sldx + (b.sx0 + b.nx - 1);
    const long dlast = ((long)(b.dz0 + b.nz - 1) * b.dldy + (b.dy0 + b.ny - 1)) * b.dldx + (b.dx0 + b.nx - 1);
    return slast < b.snelem && dlast < b.dnelem;
}

} // namespace

extern "C" {

tti_copy_plan*
tti_copy_plan_create(const tti_box_copy* boxes, int n)
{
    if (n < 0 || n > 65535 || (n > 0 && !boxes))
        return nullptr;
    auto* p = new tti_copy_plan;
    p->n = n;
    for (int i = 0; i < n; ++i) {
        if (!box_ok(boxes[i])) {
            delete p;
            return nullptr;
        }
        p->max_points = std::max(p->max_points, (long)boxes[i].nx * boxes[i].ny * boxes[i].nz);
    }
    if (n > 0) {
        if (hipMalloc((void**)&p->dev, sizeof(tti_box_copy) * n) != hipSuccess ||
            hipMemcpy(p->dev, boxes, sizeof(tti_box_copy) * n, hipMemcpyHostToDevice) != hipSuccess) {
            if (p->dev)
                (void)hipFree(p->dev);
            delete p;
            return nullptr;
        }
    }
    return p;
}

int
tti_copy_plan_run(const tti_copy_plan* plan, void* stream)
{
    if (!plan)
        return TTI_EINVAL;
    if (plan->n == 0 || plan->max_points == 0)
        return TTI_OK;
    const unsigned gx = (unsigned)std::min<long>((plan->max_points + 255) / 256, 512);
    k_copy_boxes<<<dim3(gx, (unsigned)plan->n), dim3(256), 0, (hipStream_t)stream>>>(plan->dev);
    return launch_status();
}

void
tti_copy_plan_destroy(tti_copy_plan* plan)
{
    if (!plan)
        return;
    if (plan->dev)
        (void)hipFree(plan->dev);
    delete plan;
}

} // extern "C"

// ------------------------------------------------------------------------------------------------------------------
// Box copies without allocation: the descriptor table lives in a file-scope device array instead of memory the plan
// allocates. Each GPU holds its own copy of the array at its own address, looked up once per device with
// hipGetSymbolAddress; host code passes only that address to the copies and the kernel (the array's name in host
// code is a host-side stand-in, not a device address). Every call on one GPU reuses that GPU's table, so calls that
// overlap on one GPU (different streams, or host threads) race on it.
// ------------------------------------------------------------------------------------------------------------------
static __device__ __align__(256) tti_box_copy scratch[TTI_COPY_NOALLOC_MAX];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
tti_box_copy* scratch_ptr[MAX_DEVICES] = {};

// the device address of this device's table (looked up on the first call on the device, under a lock)
int
scratch_table(tti_box_copy** out)
{
    int dev = 0;
    int rc = hip_status(hipGetDevice(&dev));
    if (rc != TTI_OK)
        return rc;
    if (dev < 0 || dev >= MAX_DEVICES)
        return TTI_EINVAL;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        rc = hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (rc != TTI_OK)
            return rc;
        scratch_ptr[dev] = static_cast<tti_box_copy*>(p);
    }
    *out = scratch_ptr[dev];
    return TTI_OK;
}

} // namespace

extern "C" {

int
tti_copy_boxes_noalloc(const tti_box_copy* boxes, int n, void* stream)
{
    if (n < 0 || n > TTI_COPY_NOALLOC_MAX || (n > 0 && !boxes))
        return TTI_EINVAL;
    long max_points = 0;
    for (int i = 0; i < n; ++i) {
        if (!box_ok(boxes[i]))
            return TTI_EINVAL;
        max_points = std::max(max_points, (long)boxes[i].nx * boxes[i].ny * boxes[i].nz);
    }
    if (n == 0 || max_points == 0)
        return TTI_OK;
    tti_box_copy* table = nullptr;
    int rc = scratch_table(&table);
    if (rc != TTI_OK)
        return rc;
    // stream-ordered after the previous call's kernel on the same stream, so back-to-back calls on one stream are safe
    rc = hip_status(
        hipMemcpyAsync(table, boxes, sizeof(tti_box_copy) * n, hipMemcpyHostToDevice, (hipStream_t)stream));
    if (rc != TTI_OK)
        return rc;
    const unsigned gx = (unsigned)std::min<long>((max_points + 255) / 256, 512);
    k_copy_boxes<<<dim3(gx, (unsigned)n), dim3(256), 0, (hipStream_t)stream>>>(table);
    return launch_status();
}

} // extern "C"
