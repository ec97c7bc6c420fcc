/*
 * C ABI of the grid-layer data movement on one GPU (conventions as in tti_hip.h).
 */
#ifndef TTI_GRID_H
#define TTI_GRID_H

#ifdef __cplusplus
extern "C" {
#endif

/* dst(dx0.., dy0.., dz0..) = src(sx0.., sy0.., sz0..) over an nx x ny x nz box; both sides strided by their own
 * leading dimensions (a contiguous buffer is a volume with ldx = nx, ldy = ny). Replaces copy_and_restride_data
 * (infrastructure/cart_volume_gpu_kernels.cu:92-121) as used by cart_volume_regular_gpu::copyData/copyFromAsync
 * (infrastructure/cart_volume_regular_gpu.cu:172-246). */
int tti_copy_box(const float* src, int sldx, int sldy, int sx0, int sy0, int sz0, float* dst, int dldx, int dldy,
                 int dx0, int dy0, int dz0, int nx, int ny, int nz, void* stream);

/* v[0..n) = value. Replaces set / set_unaligned (infrastructure/cart_volume_gpu_kernels.cu:23-90). */
int tti_fill(float* v, long n, float value, void* stream);

/* v(box) = value over an nx x ny x nz box at (x0, y0, z0) of a volume with leading dims ldx, ldy. Used by
 * cart_volume_regular_gpu::set_constant with skipHalos (infrastructure/cart_volume_regular_gpu.cu). */
int tti_fill_box(float* v, int ldx, int ldy, int x0, int y0, int z0, int nx, int ny, int nz, float value,
                 void* stream);

/* Number of floats in one halo face of width r on `axis` (0 = x, 1 = y, 2 = z): r*ny*nz, r*(nx+2r)*nz,
 * r*(nx+2r)*(ny+2r); -1 for invalid arguments. Mirrors get_size_combuf_* (infrastructure/haloManager3D_gpu.cu). */
long tti_halo_face_size(int axis, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int r);

/* Pack the interior slab of width r on side (0 low, 1 high) of axis into buf / unpack buf into the ghost slab on that
 * side. Replace pack_unpack_X/Y/Z (infrastructure/haloManager3D_gpu_kernels.cuh:286-299, :383-396, :191-204). */
int tti_halo_pack(const float* vol, int axis, int side, int ixbeg, int ixend, int iybeg, int iyend, int izbeg,
                  int izend, int ldimx, int ldimy, int r, float* buf, void* stream);
int tti_halo_unpack(float* vol, int axis, int side, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                    int ldimx, int ldimy, int r, const float* buf, void* stream);

/* Two-way ghost exchange between sub-volumes adjacent along axis (lo_vol below hi_vol), both of the same shape.
 * Replaces copy_2way_X/Y/Z (infrastructure/haloManager3D_gpu_kernels.cuh:14-31, :72-95, :133-156). */
int tti_halo_exchange_local(float* lo_vol, float* hi_vol, int axis, int ixbeg, int ixend, int iybeg, int iyend,
                            int izbeg, int izend, int ldimx, int ldimy, int r, void* stream);

/* Periodic self-exchange along axis through a device buffer of 2*tti_halo_face_size floats: the volume is its own
 * neighbour on both sides (pack both interior slabs, unpack into the opposite ghosts). */
int tti_halo_fill_periodic(float* vol, int axis, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                           int ldimx, int ldimy, int r, float* buf, void* stream);

/* One strided box copy of a batch: dst(dx0.., dy0.., dz0..) = src(sx0.., sy0.., sz0..) over nx x ny x nz points.
 * snelem/dnelem are the element counts of the two allocations (used for the bounds check at plan creation). */
typedef struct
{
    const float* src;
    float* dst;
    long snelem, dnelem;
    int sldx, sldy, sx0, sy0, sz0;
    int dldx, dldy, dx0, dy0, dz0;
    int nx, ny, nz;
} tti_box_copy;

/* A batch of box copies validated once and kept in device memory, run by one kernel launch: the local ghost
 * exchange of every sub-volume pair and field along one axis (infrastructure/haloManager3D_gpu.cu:133-205 captures
 * the same set of copies as a graph of one kernel per pair and field). Boxes of one plan must not overlap a
 * destination another box of the plan reads or writes. create allocates device memory for the table and returns NULL
 * for an invalid descriptor (out of bounds, negative size, null pointer, 2^31 points or more) or n outside
 * [0, 65535]. */
typedef struct tti_copy_plan tti_copy_plan;
tti_copy_plan* tti_copy_plan_create(const tti_box_copy* boxes, int n);
int tti_copy_plan_run(const tti_copy_plan* plan, void* stream);
void tti_copy_plan_destroy(tti_copy_plan* plan);

/* The copies of tti_copy_plan_create + tti_copy_plan_run in one call that allocates nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture, tight time loops,
 * harnesses that forbid hipMalloc). The descriptor table goes into a fixed device array of TTI_COPY_NOALLOC_MAX
 * entries (about 88 KB per GPU, held until the process ends) by a stream-ordered host-to-device copy, then one kernel
 * runs the copies. Returns TTI_EINVAL (-1) for n outside [0, TTI_COPY_NOALLOC_MAX] or an invalid descriptor (the same
 * checks as tti_copy_plan_create), a negative HIP status otherwise on failure. Every call on one GPU shares that GPU's
 * table: calls on one stream run in order, calls that overlap on one GPU (other streams, other host threads) race. */
#define TTI_COPY_NOALLOC_MAX 1024
int tti_copy_boxes_noalloc(const tti_box_copy* boxes, int n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TTI_GRID_H */
