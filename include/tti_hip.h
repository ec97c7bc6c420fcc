/*
 * C ABI of the MI355X (gfx950) TTI finite-difference hot path.
 *
 * Every entry point takes device pointers at the base of a volume (ghosts included), inclusive index ranges in
 * array coordinates, leading dimensions, and an opaque stream (hipStream_t, NULL = the null stream). Each returns
 * 0 on success, -1 for an invalid argument, -2 for an unsupported FD order, and -1000 - hipError_t when the HIP
 * runtime reported an error. No call allocates device memory unless its comment says so.
 *
 * Volumes are FP32, x fastest: idx = (iz*ldimy + iy)*ldimx + ix (utils/volume_index.h:22). Stencil entry points need
 * a ghost zone of at least 4 points on every side of the interior range; the x and y extents are checked against
 * ldimx/ldimy, the z extent is the caller's contract (as in the reference launchers, which take no ldimz).
 *
 * Names and argument meaning follow the reference launchers they replace (file:line in each comment); bool
 * arguments become int and cudaStream_t becomes void*.
 */
#ifndef TTI_HIP_H
#define TTI_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- type2 (variable-density TTI) operators ------------------------------------------------------------------- */

/* d_a f = Model11 * D-_a f for f in {field1, field2}, a in {x, y, z}; writes the interior of the six outputs.
 * Replaces kernel3_gpu_kernels::launch_fwd_kernel2_drv (kernels/kernel_type2_fwd_drv.h:7-11). order 8 or 16. */
int launch_fwd_kernel2_drv(const float* field1, const float* field2, const float* Model11, float* dfield1dx,
                           float* dfield1dy, float* dfield1dz, float* dfield2dx, float* dfield2dy, float* dfield2dz,
                           int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, float invdx, float invdy,
                           float invdz, int ldimx, int ldimy, int order, void* stream);

/* field3 += dt*rhs3, field4 += dt*rhs4 from the six derivatives, the rotated mixed derivatives of field1/field2 and
 * the models (irho = Model11 = 1/rho). Replaces launch_update_rho_fwd_main_loop_2_1 (kernels/kernel_type2_fwd.h:6-14).
 * The PML pointers are accepted and ignored; sponge_active or any pml flag returns -1 (the reference instantiates only
 * the interior form, kernels/kernel_type2_fwd.cu:641,:657). Sponge pointers are float2 arrays {A, B}. */
int launch_update_rho_fwd_main_loop_2_1(float* field3, float* field4, const float* field1_dx, const float* field2_dx,
                                        const float* field1_dy, const float* field2_dy, const float* field1_dz,
                                        const float* field2_dz, const float* field1, const float* field2,
                                        const float* irho, const float* model1, const float* model2,
                                        const float* model3, const float* model4, const float* rx, const float* ry,
                                        const float* rz, float* pml_field1_xx, float* pml_field2_xx,
                                        float* pml_field1_yy, float* pml_field2_yy, float* pml_field1_zz,
                                        float* pml_field2_zz, const void* pml_ab_x, const void* pml_ab_y,
                                        const void* pml_ab_z, const void* sponge_ab_xx, const void* sponge_ab_yy,
                                        const void* sponge_ab_zz, int ixbeg, int ixend, int iybeg, int iyend,
                                        int izbeg, int izend, int ldimx, int ldimy, float dt, double dx, double dy,
                                        double dz, int sponge_active, int order, int pml_x, int pml_y, int pml_z,
                                        void* stream);

/* Adjoint derivatives: S = rho*(lap + r_a^2 hz), d_a = Model11*D-_a S, rhs = sum X_ab(2 r_a r_b hz).
 * Replaces adj_kernel3_derivatives::launch_adj_kernel2_drv (kernels/kernel_type2_grad_snap.h:7-12).
 * invdxy/invdyz/invdzx are 1/(4 da db) as the reference passes them (kernels/kernel_type2.cu:236-237);
 * invdxx/invdyy/invdzz and dt are accepted and unused, as in the reference kernel. */
int launch_adj_kernel2_drv(float* dfield1dx, float* dfield2dx, float* dfield1dy, float* dfield2dy, float* dfield1dz,
                           float* dfield2dz, float* field3_rhs, float* field4_rhs, const float* field1,
                           const float* field2, const float* Model11, const float* model1, const float* model2,
                           const float* model3, const float* model4, const float* model5, const float* model6,
                           const float* model7, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                           float invdxx, float invdyy, float invdzz, float invdxy, float invdyz, float invdzx,
                           int ldimx, int ldimy, int order, float dt, float invdx, float invdy, float invdz,
                           void* stream);

/* field3 += dt*(sum_a D+_a d_a + field3_rhs) (or A*field3 + B*dt*rhs when sponge_active), field4 likewise.
 * Replaces kernel3_gpu_kernels::launch_update_adj_kernel_main_3 (kernels/kernel_type2_adj.h:8-14). */
int launch_update_adj_kernel_main_3(float* field3, float* field4, const float* field1_dx, const float* field2_dx,
                                    const float* field1_dy, const float* field2_dy, const float* field1_dz,
                                    const float* field2_dz, const float* field3_rhs, const float* field4_rhs,
                                    const void* ab_xx, const void* ab_yy, const void* ab_zz, int ixbeg, int ixend,
                                    int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy, float dt,
                                    double dx, double dy, double dz, int pmlx, int pmly, int pmlz, int order,
                                    int sponge_active, void* stream);

/* Density gradient grad += -b * sum [S D+ Dfwd + (D- S) Dfwd]; grad is NOT shifted by the interior offset (its
 * element (0,0,0) is the interior origin), exactly as the reference kernel indexes it.
 * Replaces launch_update_rho_gradient_rho (kernels/kernel_type2_grad.h:22-32); both values of simple_gradient
 * compute the same quantity here. */
int launch_update_rho_gradient_rho(float* grad, const float* fwd_field1_dx, const float* fwd_field2_dx,
                                   const float* fwd_field1_dy, const float* fwd_field2_dy, const float* fwd_field1_dz,
                                   const float* fwd_field2_dz, const float* field1, const float* field2,
                                   const float* irho, const float* model1, const float* model2, const float* model3,
                                   const float* model4, const float* rx, const float* ry, const float* rz, int ixbeg,
                                   int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy,
                                   int grad_ldimx, int grad_ldimy, double dx, double dy, double dz, int order,
                                   int simple_gradient, void* stream);

/* ---- time update, free surface, imaging -------------------------------------------------------------------------- */

/* field1 = A*field1 + B*dt*field3, field2 likewise, A/B = min over the axes' sponge pairs (float2 {A,B} indexed by
 * array index). Replaces kernel_base::kernel_loop2_inner / kernel_loop2_inner_kernel (kernels/kernel_base.cu:82-140). */
int kernel_loop2_inner(float* field1, float* field2, const float* field3, const float* field4, const void* ABx,
                       const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg,
                       int izend, int ldimx, int ldimy, float dt, void* stream);

/* ---- viscoacoustic memory variables (nmech = 1..3 mechanisms; the reference uses 3, kernels/kernel_base.h:27) ----
 * Model8[m], Model9[m]: memory variables of field1 / field2 (device volumes of the fields' layout), Model10[m]: alpha_m,
 * wsi[m]: omega_m (host array). The pointer arrays are host arrays of nmech device pointers. Interior points only;
 * ABx/ABy/ABz as in kernel_loop2_inner, or all NULL for A = B = 1. */

/* field1 = A*field1 + B*dt*(field3 - sum_m wsi[m]*Model8[m]), field2 likewise with field4, Model9 (in place of loop2).
 * Replaces kernel_base::kernel_loop1_inner / kernel_loop1_inner_kernel (kernels/kernel_base.cu:189-272). */
int kernel_loop1_inner(float* field1, float* field2, const float* field3, const float* field4,
                       const float* const* Model8, const float* const* Model9, const float* wsi, int nmech,
                       const void* ABx, const void* ABy, const void* ABz, int ixbeg, int ixend, int iybeg, int iyend,
                       int izbeg, int izend, int ldimx, int ldimy, float dt, void* stream);

/* Model8[m] += dt*(Model10[m]*field3 - wsi[m]*Model8[m]), Model9[m] likewise with field4.
 * Replaces kernel_base::kernel_loop3_inner / kernel_loop3_inner_kernel (kernels/kernel_base.cu:274-364) without the
 * Born fields (kernel_loop3_inner_born below has them). */
int kernel_loop3_inner(const float* field3, const float* field4, float* const* Model8, float* const* Model9,
                       const float* const* Model10, const float* wsi, int nmech, int ixbeg, int ixend, int iybeg,
                       int iyend, int izbeg, int izend, int ldimx, int ldimy, float dt, void* stream);

/* kernel_loop3_inner on two sets of memory variables in one pass: Model8/Model9 from field3/field4 and the Born
 * memory variables Model81/Model91 from the Born fields field31/field41, with the same Model10[m] (read once per point)
 * and wsi[m]; each set's result equals kernel_loop3_inner on that set alone. Replaces the update_born branch of
 * kernel_base::kernel_loop3_inner / kernel_loop3_inner_kernel<nmech, true> (kernels/kernel_base.cu:274-364). */
int kernel_loop3_inner_born(const float* field3, const float* field4, const float* field31, const float* field41,
                            float* const* Model8, float* const* Model9, float* const* Model81, float* const* Model91,
                            const float* const* Model10, const float* wsi, int nmech, int ixbeg, int ixend, int iybeg,
                            int iyend, int izbeg, int izend, int ldimx, int ldimy, float dt, void* stream);

/* field1 = A*field1 + B*dt*(field3 + sum_m Model10[m]*Model8[m]), field2 likewise (adjoint, in place of loop2).
 * Replaces kernel_base::adj_kernel_loop1_inner (kernels/kernel_base.cu:366-431). */
int adj_kernel_loop1_inner(float* field1, float* field2, const float* field3, const float* field4,
                           const float* const* Model8, const float* const* Model9, const float* const* Model10,
                           int nmech, const void* ABx, const void* ABy, const void* ABz, int ixbeg, int ixend,
                           int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy, float dt, void* stream);

/* Model8[m] += dt*wsi[m]*(-field3 - Model8[m]), Model9[m] likewise with field4 (adjoint).
 * Replaces kernel_base::adj_kernel_loop3_inner (kernels/kernel_base.cu:433-492). */
int adj_kernel_loop3_inner(const float* field3, const float* field4, float* const* Model8, float* const* Model9,
                           const float* wsi, int nmech, int ixbeg, int ixend, int iybeg, int iyend, int izbeg,
                           int izend, int ldimx, int ldimy, float dt, void* stream);

/* Free-surface mirror of field1/field2 at plane izbeg over whole xy planes.
 * Replaces kernel_base::kernel_loop4_inner (kernels/kernel_base.cu:142-187). */
int kernel_loop4_inner(float* field1, float* field2, int ldimx, int ldimy, int izbeg, void* stream);

/* Free-surface mirror of the six derivatives (x, y odd; z even about izbeg + 1/2).
 * Replaces kernel_type2::kernel_loop4_derivative_impl / kernel_loop4_derivatives (kernels/kernel_type2_small.cu:5-40,
 * kernels/kernel_type2.cu:279-293). */
int kernel_loop4_derivatives(float* dfield1dx, float* dfield1dy, float* dfield1dz, float* dfield2dx, float* dfield2dy,
                             float* dfield2dz, int ldimx, int ldimy, int izbeg, void* stream);

/* grad(no halo) += adj3*snap3 + adj4*snap4; snap and grad share the no-halo layout (snap_ldimx, snap_ldimy).
 * Replaces update_Vp_gradient / kernel_type2::update_grad_loop_1 (kernels/kernel_type2_small.cu:42-59,
 * kernels/kernel_type2.cu:306-330) and launch_update_grad_loop_kernel (kernels/kernel_type1_gpu_grad.h:4-7). */
int update_Vp_gradient(const float* snap_field3, const float* snap_field4, float* grad_Vp, const float* adj_field3,
                       const float* adj_field4, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                       int snap_ldimx, int snap_ldimy, int ldimx, int ldimy, void* stream);

/* grad *= -dt_xcorr*2/sqrt(model3). Replaces finalize_Vp_gradient (kernels/kernel_type2_small.cu:61-76). */
int finalize_Vp_gradient(const float* model3, float* grad_Vp, int ixbeg, int ixend, int iybeg, int iyend, int izbeg,
                         int izend, int grad_ldimx, int grad_ldimy, int ldimx, int ldimy, float dt_xcorr,
                         void* stream);

/* Models m1..m7 and Model11 = 1/rho from per-point (vp, eps, delta, vs, theta, phi, rho) over n elements.
 * Replaces simulator2SetUp::init_cijs / init_cijs_kernel (simulator2/simulator2SetUp.cu:10-29), which takes
 * constants and never sets Model11. */
int init_cijs(long n, const float* vp, const float* eps, const float* delta, const float* vs, const float* theta,
              const float* phi, const float* rho, float* model1, float* model2, float* model3, float* model4,
              float* model5, float* model6, float* model7, float* Model11, void* stream);

/* pressure = -(2*field1 + field2)/3 over n elements. Replaces compute_pressure_kernel / computefield5
 * (simulator2/simulator2SetUp.cu:31-42, :842-857). */
int compute_pressure(long n, const float* field1, const float* field2, float* pressure, void* stream);


/* ---- type1 (constant-density TTI) operators -------------------------------------------------------------------- */

/* Constant-density TTI forward: F_aa central second derivatives (1/da^2), F_ab rotated mixed (1/(4 da db)), same V/H
 * model combination as the type2 update without rho. calc_snap = 1 writes rhs into field3/field4 with the no-halo
 * layout (ldimx_out, ldimy_out); otherwise field3 += dt*rhs, or field3 = A*field3 + B*dt*rhs when use_sponge (or
 * simple, whose reference kernel always applies the sponge form). ldimz_in/out are accepted and unused.
 * Replaces fwd_main_loop_2_1_inner_gpu::launch (kernels/kernel_type1_gpu_fwd.h:6-10). The x-z mixed stencil is the
 * correct one of the optimised reference kernel for both values of `simple`. */
int fwd_main_loop_2_1_inner_gpu_launch(float* field3, float* field4, const float* field1, const float* field2,
                                       const float* model1, const float* model2, const float* model3,
                                       const float* model4, const float* model5, const float* model6,
                                       const float* model7, const void* ABx, const void* ABy, const void* ABz,
                                       int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx_in,
                                       int ldimy_in, int ldimz_in, int ldimx_out, int ldimy_out, int ldimz_out,
                                       float dt, double dx, double dy, double dz, int order, void* stream,
                                       int calc_snap, int use_sponge, int simple);

/* Constant-density TTI adjoint: field3 += dt*(sum_a central2_a(lap + r_a^2 hz) + sum X_ab(2 r_a r_b hz)) (sponge
 * form with use_sponge or simple). Replaces adj_main_loop_2_inner_gpu::launch (kernels/kernel_type1_gpu_adj.h:6-10). */
int adj_main_loop_2_inner_gpu_launch(float* field3, float* field4, const float* field1, const float* field2,
                                     const float* model1, const float* model2, const float* model3,
                                     const float* model4, const float* model5, const float* model6,
                                     const float* model7, const void* ABx, const void* ABy, const void* ABz, int ixbeg,
                                     int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy,
                                     int ldimz, float dt, double dx, double dy, double dz, int order, void* stream,
                                     int use_sponge, int simple);

/* Imaging condition of the type1 family (same formula as update_Vp_gradient); the reference's volume_index
 * arguments become the explicit leading dimensions. Replaces launch_update_grad_loop_kernel
 * (kernels/kernel_type1_gpu_grad.h:4-7). */
int launch_update_grad_loop_kernel(const float* snap_x, const float* snap_z, float* grad_Vp, const float* adj_x,
                                   const float* adj_z, int ixbeg, int ixend, int iybeg, int iyend, int izbeg,
                                   int izend, int snap_ldimx, int snap_ldimy, int ldimx, int ldimy, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TTI_HIP_H */
