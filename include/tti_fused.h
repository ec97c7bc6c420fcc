/*
 * C ABI of the fused gfx950 time steps (one kernel per forward step, one per adjoint step), the receiver/source
 * helpers and the compressed snapshot format. Conventions (pointers, ranges, status codes, stream) are those of
 * tti_hip.h.
 *
 * These entry points have no one-to-one counterpart in the reference: each replaces a sequence of reference calls
 * made by simulator2SetUp::execute (simulator2/simulator2SetUp.cu:530-779) and computes the same values.
 */
#ifndef TTI_FUSED_H
#define TTI_FUSED_H

#ifdef __cplusplus
extern "C" {
#endif

/* Snapshot storage of the forward rates u (field3) and w (field4) for the imaging condition. Arrays cover the
 * interior only, shape (nz, ldimy, ldimx) with ldimx a multiple of 64 and >= nx, ldimy >= ny.
 *   TTI_SNAP_FP32: float arrays.
 *   TTI_SNAP_FP16: IEEE half arrays plus one float scale per 64-point x-row segment (scale3/scale4 of shape
 *                  (nz, ldimy, ldimx/64)); value = half * scale, scale a power of two chosen from the segment max.
 *   TTI_SNAP_BF16: bfloat16 arrays, no scale. */
#define TTI_SNAP_NONE 0
#define TTI_SNAP_FP32 1
#define TTI_SNAP_FP16 2
#define TTI_SNAP_BF16 3

typedef struct
{
    int mode;
    void* field3;
    void* field4;
    float* scale3;
    float* scale4;
    int ldimx;
    int ldimy;
    /* Optional (NULL: not stored): sigma_H (field1) and sigma_V (field2) of the step in the same encoding and layout,
     * for the density image. A forward step stores the sigma it read, the one its u, w update used (the snapshot
     * pair (u, sigma) of simulator2/simulator2SetUp.cu:606-645). */
    void* field1;
    void* field2;
    float* scale1;
    float* scale2;
} tti_snapshot;

/* One forward time step in a single sweep. Inputs: field1/field2 = sigma after the previous time update (read with
 * ghosts), field3/field4 = u, w (updated in place). Computes
 *   u += dt*L(sigma)  [as launch_fwd_kernel2_drv followed by launch_update_rho_fwd_main_loop_2_1]
 *   u += src_value, w += src_value at linear index src_index (none if < 0)
 *   field1_out = A*field1 + B*dt*u, field2_out likewise [kernel_loop2_inner of the next step]
 * Sponge arrays ABx/ABy/ABz are float2 {A,B} per array index, or all NULL for A = B = 1.
 * field*_out must not alias field1/field2. When snap is non-NULL and snap->mode != 0 the updated u, w of every interior
 * plane are stored into it. free_surface != 0 applies the stress-free surface at plane izbeg as the reference's
 * mirrors do (kernel_loop4_derivatives on the z derivatives inside the sweep, then kernel_loop4_inner on field*_out;
 * field1/field2 must already be mirrored, as the previous step or kernel_loop4_inner leaves them). zblock = z planes
 * per thread block (0 = default 128); strip = block schedule: the x-y tiles of a z chunk are visited in vertical strips
 * of `strip` tiles of 64 columns (0 = full rows). zblock and strip change the speed, never the result.
 * Replaces the forward sequence of simulator2SetUp::execute (simulator2/simulator2SetUp.cu:536-599) on one volume. */
int tti_fused_fwd_step(const float* field1, const float* field2, float* field1_out, float* field2_out, float* field3,
                       float* field4, const float* model1, const float* model2, const float* model3,
                       const float* model4, const float* rx, const float* ry, const float* rz, const float* Model11,
                       const void* ABx, const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend,
                       int izbeg, int izend, int ldimx, int ldimy, float dt, double dx, double dy, double dz,
                       long src_index, float src_value, const tti_snapshot* snap, int free_surface, int zblock,
                       int strip, void* stream);

/* One adjoint time step in a single sweep, in the reference order (adjoint derivatives, update, time update):
 *   field3 += dt*L^T(field1, field2)   (sponge form A*field3 + B*dt*L^T(...) when ABx is given)
 *   field1_out = A*field1 + B*dt*field3, field2_out likewise
 *   grad(interior, FP64, shape nz x ny x nx) += field3*snap3 + field4*snap4  when snap and grad are non-NULL
 * free_surface, zblock and strip as for tti_fused_fwd_step (the mirrors of simulator2/simulator2SetUp.cu:690-730).
 * Replaces adj_main_loop + adj_main_loop_2 + kernel_loop2 + update_grad_loop_1 of simulator2SetUp::execute
 * (simulator2/simulator2SetUp.cu:675-745). */
int tti_fused_adj_step(const float* field1, const float* field2, float* field1_out, float* field2_out, float* field3,
                       float* field4, const float* model1, const float* model2, const float* model3,
                       const float* model4, const float* rx, const float* ry, const float* rz, const float* Model11,
                       const void* ABx, const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend,
                       int izbeg, int izend, int ldimx, int ldimy, float dt, double dx, double dy, double dz,
                       const tti_snapshot* snap, double* grad, int free_surface, int zblock, int strip,
                       void* stream);

/* Density image of one snapshot step: the composition update_grad_loop_2 of simulator2SetUp::execute
 * (simulator2/simulator2SetUp.cu:745-767: snaps_loop, then launch_update_rho_gradient_rho), accumulated in FP64:
 *   Dfwd = Model11 * D- sigma_snap                       (sigma_snap = snap->field1/field2, decoded to FP32)
 *   grad_rho(interior, FP64, nz x ny x nx) += -b * sum_{f,a} [S_{f,a} D+_a Dfwd_{f,a} + (D-_a S_{f,a}) Dfwd_{f,a}]
 * with S from the adjoint sigma field1/field2 (after the adjoint step and any residual injection) and the models.
 * work = 8 FP32 volumes of the volumes' layout (ldimx * ldimy * (izend + 5) floats each, back to back) that the caller
 * zeroes once and keeps: only their interiors are written, so their ghosts stay zero. Order 8. */
int tti_rho_image(const float* field1, const float* field2, const float* model1, const float* model2,
                  const float* model3, const float* model4, const float* rx, const float* ry, const float* rz,
                  const float* Model11, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx,
                  int ldimy, double dx, double dy, double dz, const tti_snapshot* snap, float* work, double* grad_rho,
                  void* stream);

/* out[k] = -(2*field1 + field2)/3 at linear indices index[k] (pressure at receivers; the pressure definition of
 * simulator2/simulator2SetUp.cu:31-42). */
int tti_gather_pressure(const float* field1, const float* field2, const long* index, int n, float* out, void* stream);

/* Transpose of tti_gather_pressure applied after tti_fused_adj_step: field3 += -2/3*r, field4 += -1/3*r at index[k]
 * and field*_out += B*dt*(same increment); ijk[3k..3k+2] are the array coordinates of index[k] (distinct points).
 * With snap/grad non-NULL the increment's imaging-condition term is added to grad as well. */
int tti_inject_residual(float* field3, float* field4, float* field1_out, float* field2_out, const long* index,
                        const int* ijk, int n, const float* residual, float dt, const void* ABx, const void* ABy,
                        const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                        const tti_snapshot* snap, double* grad, void* stream);

/* Decodes field3 (which = 3), field4 (4), field1 (1) or field2 (2) of a snapshot into a dense FP32 (nz, ny, nx)
 * array. */
int tti_snapshot_decode(const tti_snapshot* snap, int which, int nx, int ny, int nz, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TTI_FUSED_H */
