/* C ABI of the staggered-grid elastic TTI ("eTTI") operators of the simulator3 driver, gfx950 build.
 *
 * Volumes: FP32 device pointers at the volume base, x fastest, idx = (iz*ldimy + iy)*ldimx + ix, 4-point ghosts; the
 * interior is [4, 4+n) on each axis and the caller's allocation holds nz + 8 planes (the z extent is not passed, as
 * in the reference launchers). Only interior points are written unless stated otherwise. Every entry returns 0 on
 * success, -1 for bad arguments, -1000 - hipError_t for a HIP failure; `stream` is a hipStream_t (NULL = default).
 *
 * Coefficients: the stencils read two process-wide 4-entry tables, df (staggered first derivative,
 * D+ g(i) = sum_k df[k] (g(i+k+1) - g(i-k)), D- g(i) = sum_k df[k] (g(i+k) - g(i-k-1))) and si (diagonal interpolation,
 * Mix g = 0.5 sum_n si[n-1] sum_{p,q in {n,1-n}} g(i_a + p - s_a, i_b + q - s_b)). They are zero until
 * etti_init_constant sets them, exactly as the reference runs (it never calls its init_constant).
 */
#ifndef TTI_ETTI_H
#define TTI_ETTI_H

#ifdef __cplusplus
extern "C" {
#endif

/* Set the coefficient tables (either may be NULL to keep it). Replaces init_constant
 * (simulator3/simulator3SetUp.cu:28-34), which would upload {1.01, 2.01, 3.01, 4.01} to both. */
int etti_init_constant(const float* df8, const float* s8);
int etti_get_constant(float* df8, float* s8);

/* f1 += 2 M22(x-1) (1.01 D-x f4 + 2.01 D+y f7 + zprime[z] D+z f8), f2 += 2 M22(y-1) (1.01 D+x f7 + 2.01 D-y f5 +
 * zprime[z] D+z f9), f3 += 2 M22(z-1) (1.01 D+x f8 + 2.01 D+y f9 + zprime[z] D-z f6). milliseconds_v, when not NULL,
 * accumulates the kernel time (the call then waits for it), as the reference's float& argument does.
 * Replaces vPlainLoop (simulator3/simulator3SetUp.cu:1106-1125) / vPlainKernel (:820-975). */
int vPlainLoop(float* f1, float* f2, float* f3, const float* f4, const float* f5, const float* f6, const float* f7,
               const float* f8, const float* f9, const float* M22, const float* zprime, int nx, int ny, int nz,
               int ldimx, int ldimy, float* milliseconds_v, void* stream);

/* d4 = 1.01 D+x f1, d5 = 2.02 D-y f1 + 1.01 D-x f2, d6 = zp D-z f1 + 1.01 D-x f3, d7 = 2.02 D+y f2,
 * d8 = zp D-z f2 + 2.02 D-y f3, d9 = zp D+z f3 (zp = zprime[z]).
 * Replaces vgPlainLoop (simulator3/simulator3SetUp.cu:1363-1377) / vgPlainKernel (:1218-1361). */
int vgPlainLoop(const float* f1, const float* f2, const float* f3, float* d4, float* d5, float* d6, float* d7,
                float* d8, float* d9, const float* zprime, int nx, int ny, int nz, int ldimx, int ldimy,
                float* milliseconds_g, void* stream);

/* drvxz = Mix_xz(input; xs, zs), drvyz = Mix_yz(input; ys, zs), shifts 0 or 1.
 * Replaces the template kernel precompute_xz_yz<XS,YS,ZS> (simulator3/simulator3SetUp.cu:47-176); every output plane
 * weights each input plane by its offset (the reference's first outputs of each z block are weighted by the priming
 * step instead, :120-125, which its zero coefficients hide). */
int precompute_xz_yz(int xs, int ys, int zs, const float* input, float* drvxz, float* drvyz, int nx, int ny, int nz,
                     int ldimx, int ldimy, void* stream);

/* The six precompute_xz_yz launches into f10..f21, then the stress update of sPlainKernel_launch; milliseconds_s
 * times the update only. f1_1, f1_2, f1_3, f2_2, f2_3, f3_3 are the strains d4, d5, d6, d7, d8, d9.
 * Replaces sPlainLoop (simulator3/simulator3SetUp.cu:1128-1216). */
int sPlainLoop(float* f4, float* f5, float* f6, float* f7, float* f8, float* f9, const float* M1, const float* M2,
               const float* M3, const float* M4, const float* M5, const float* M6, const float* M7, const float* M8,
               const float* M9, const float* M10, const float* M11, const float* M12, const float* M13,
               const float* M14, const float* M15, const float* M16, const float* M17, const float* M18,
               const float* M19, const float* M20, const float* M21, const float* f1_1, const float* f1_2,
               const float* f1_3, const float* f2_2, const float* f2_3, const float* f3_3, float* f10, float* f11,
               float* f12, float* f13, float* f14, float* f15, float* f16, float* f17, float* f18, float* f19,
               float* f20, float* f21, int nx, int ny, int nz, int ldimx, int ldimy, float* milliseconds_s,
               void* stream);

/* Stress update alone: f4..f9 += C(M1..M21) applied to the strains, with the six in-plane mixes computed from
 * d4..d9 and the twelve x-z / y-z mixes read from f10..f21. Arrays are host arrays of device pointers (6, 21, 6, 12).
 * Replaces the sPlainKernel launch (simulator3/simulator3SetUp.cu:1193-1214; kernel :977-1104). */
int sPlainKernel_launch(float* const* f4_f9, const float* const* M1_M21, const float* const* d4_d9,
                        const float* const* f10_f21, int nx, int ny, int nz, int ldimx, int ldimy, void* stream);

/* Free surface over whole x-y planes (stride = ldimx*ldimy points): planes 0..izbeg, zero at izbeg, f(iz) =
 * -f(2 izbeg - iz) above it. Replace mirror_s / mirror_s_kernel and mirror_v / mirror_v_kernel
 * (simulator3/simulator3SetUp.cu:1379-1480). */
int mirror_s_inner(float* f4, float* f5, float* f6, float* f7, float* f8, float* f9, long stride, int izbeg,
                   void* stream);
int mirror_v_inner(float* f1, float* f2, float* f3, long stride, int izbeg, void* stream);

/* g1 += a_f22 snap_f22 + a_f23 snap_f23 over the inclusive ranges of the a volumes; g1 and the snapshots have no
 * ghosts (element 0 is the range origin) and leading dims snap_ldimx, snap_ldimy.
 * Replaces update_g_v / update_gv (simulator3/simulator3SetUp.cu:1482-1538). */
int update_gv_inner(const float* snap_f22, const float* snap_f23, float* g1, const float* a_f22, const float* a_f23,
                    int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int snap_ldimx, int snap_ldimy,
                    int ldimx, int ldimy, void* stream);

/* Density imaging condition g += -m18 sum[...] with g not shifted by the interior offset; the same operator as
 * launch_update_rho_gradient_rho with (f24, f25, m18) in the (field1, field2, Model11) slots.
 * Replaces launch_update_g_r (simulator3/simulator3SetUp.cu:2037-2089; kernels :1841-2004, :2130-2179). */
int launch_update_g_r(float* g, const float* f_f24_dx, const float* f_f25_dx, const float* f_f24_dy,
                      const float* f_f25_dy, const float* f_f24_dz, const float* f_f25_dz, const float* f24,
                      const float* f25, const float* m18, const float* m1, const float* m2, const float* m3,
                      const float* m4, const float* rx, const float* ry, const float* rz, int ixbeg, int ixend,
                      int iybeg, int iyend, int izbeg, int izend, int ldimx, int ldimy, int g_ldimx, int g_ldimy,
                      double dx, double dy, double dz, int order, int simple_kernel, void* stream);

/* Synthetic model of init_cijs_kernel (simulator3/simulator3SetUp.cu:230-271): M1..M22 and d4..d9 (host arrays of 22
 * and 6 device pointers; d4_d9 may be NULL to fill the models only) get cos/sin of weighted index sums for x in
 * [ixbeg, ixend] and every y < ldimy, z < ldimz; zprime[z] = cos(6 z) for z < ldimz. */
int etti_init_cijs(float* const* M1_M22, float* const* d4_d9, float* zprime, int ixbeg, int ixend, int ldimx,
                   int ldimy, int ldimz, void* stream);

/* p = -(2 f4 + 3 f5 + 4 f6 + 5 f7 + 6 f8 + 7 f9) / 27 over n points (compute_p_kernel, simulator3/simulator3SetUp.cu:
 * 273-287, launched over the whole allocation by computeP, :2691-2708). */
int etti_compute_p(float* p, const float* f4, const float* f5, const float* f6, const float* f7, const float* f8,
                   const float* f9, long n, void* stream);

/* f4..f9 at one index = p * (1, 2, 3, 4, 5, 6) (init_src_kernel, simulator3/simulator3SetUp.cu:289-300). */
int etti_init_src(float* f4, float* f5, float* f6, float* f7, float* f8, float* f9, long index, float p,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TTI_ETTI_H */
