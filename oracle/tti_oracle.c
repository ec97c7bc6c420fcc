/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle for the TTI finite-difference hot path.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and only as the
 * reference that GPU results are compared against (or, in bench.py, as the timed CPU baseline). The product path
 * (miniappsforhpe_amd) never links or calls it and fails loudly when its HIP library is missing.
 *
 * This file restates, in plain C, the arithmetic of the reference kernels of
 * knsuresh2806/MiniAppsForHPE (simulatorsForNvidia/...). Every function cites the reference file:line it follows.
 * It is compiled twice:
 *   liboracle64.so : ORC_REAL = double  -> the FP64 oracle used by the parity tests
 *   liboracle32.so : ORC_REAL = float   -> the same formulas in FP32 with OpenMP, the CPU baseline in bench.py
 *
 * Conventions (same as the reference):
 *   volumes are x-fastest: idx = (iz*ldy + iy)*ldx + ix   (utils/volume_index.h:22)
 *   index ranges [ibeg, iend] are inclusive and in array coordinates (ghosts included)
 *   coefficients are the reference's FP32 constants (kernels/fd_coef.h:467-470, :517-521), widened exactly
 *
 * Pinning (see tests/test_oracle_math.py): closed-form coefficient identities, plane-wave symbols of every
 * derivative operator, TTI invariants, exact discrete transposes (dot-product tests) and brute-force index checks.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;

/* float2 as laid out by the reference sponge holders (utils/sponge_coef3d_gpu.h:17-29): {A, B} per index */
typedef struct
{
    float a, b;
} orc_ab;

#define IDX(ix, iy, iz) (((long)(iz) * ldy + (long)(iy)) * ldx + (long)(ix))

/* ---------------------------------------------------------------------------------------------------------------
 * FD coefficients, _FD_COEF_LEAST_SQUARE, as FP32 constants (converted to float first, as the reference's
 * `const float` tables are, then widened):
 *   order 8  : first staggered c1..c4 kernels/fd_coef.h:467-470; central second e0..e4 :517-521
 *   order 16 : first staggered c1..c8 kernels/fd_coef.h:479-486; central second e0..e8 :535-543
 *   order 12 : first staggered c1..c6 of stg_d1<12> (kernels/fd_coefficients.h:16-60), which only the density
 *              gradient takes (kernels/kernel_type2_grad.cu:612-623); no central set
 * The active order (8 by default) is set by orc_set_order; every operator below uses radius ORAD = order / 2.
 * ------------------------------------------------------------------------------------------------------------- */
static const float C8F[9] = { 0.0f, 1.19628906250000e+00f, -7.97526041666667e-02f, 9.57031250000000e-03f,
                              -6.97544642857143e-04f };
static const float E8F[9] = { -0.2903117E+1f, 0.1645351E+1f, -0.2236606E+0f, 0.3265851E-1f, -0.2790073E-2f };
static const float C16F[9] = { 0.0f, 1.254298E+00f, -1.234902E-01f, 3.470321E-02f, -1.201822E-02f,
                               4.188332E-03f, -1.310232E-03f, 3.233353E-04f, -4.675068E-05f };
static const float E16F[9] = { -0.3162230E+1f, 0.1875943E+1f, -0.3859840E+0f, 0.1228261E+0f, -0.4233006E-1f,
                               0.1379639E-1f, -0.3846673E-2f, 0.8026637E-3f, -0.9257725E-4f };
static const float C12F[9] = { 0.0f, 1.2403120e+0f, -1.1141390e-1f, 2.5759060e-2f, -6.4207590e-3f,
                               1.3095870e-3f, -1.5512770e-4f };
static const float ZEROF[9] = { 0.0f };

static const float* CACT = C8F;
static const float* EACT = E8F;
static int ORAD = 4;

/* Selects the FD order of every operator (8, 12 or 16); returns 0, or -1 for another order. Order 12 has no central
 * set (only the staggered operators are defined for it). */
int
orc_set_order(int order)
{
    if (order == 8) {
        CACT = C8F;
        EACT = E8F;
    } else if (order == 16) {
        CACT = C16F;
        EACT = E16F;
    } else if (order == 12) {
        CACT = C12F;
        EACT = ZEROF;
    } else {
        return -1;
    }
    ORAD = order / 2;
    return 0;
}

#define C(n) ((real)CACT[n])
#define E(n) ((real)EACT[n])

int
orc_real_size(void)
{
    return (int)sizeof(real);
}

/* the active coefficient sets, c[0..8] (c[0] = 0) and e[0..8], zero beyond the radius */
void
orc_coefs(double* c, double* e)
{
    for (int i = 0; i < 9; ++i) {
        c[i] = i <= ORAD ? (double)CACT[i] : 0.0;
        e[i] = i <= ORAD ? (double)EACT[i] : 0.0;
    }
}

/* kernel_utils::compute_fd_const (kernels/kernel_utils.h:83-97): inverses are computed in double and stored as
 * float by the reference; the oracle takes them already rounded by the caller (tests pass float-rounded values). */

/* ---------------------------------------------------------------------------------------------------------------
 * a2: forward staggered derivatives d_a f = b * D-_a f, f in {field1, field2}, a in {x, y, z}
 *     D-_a g(i) = invd * sum_n c_n [ g(i+n-1) - g(i-n) ]   (value at i-1/2), b taken at the integer point i.
 *     kernels/kernel_type2_fwd_drv.cu:50-76 (z), :78-110 (x, y), :112-130 (multiply by Model11), launcher :368-409.
 *     Writes the interior [ixbeg..ixend] x [iybeg..iyend] x [izbeg..izend] only.
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_fwd_drv(const real* f1, const real* f2, const real* b, real* d1x, real* d1y, real* d1z, real* d2x, real* d2y,
            real* d2z, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, real invdx, real invdy,
            real invdz, long ldx, long ldy)
{
    const long sy = ldx, sz = ldx * ldy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                real ax1 = 0, ay1 = 0, az1 = 0, ax2 = 0, ay2 = 0, az2 = 0;
                for (int n = 1; n <= ORAD; ++n) {
                    ax1 += C(n) * (f1[i + (n - 1)] - f1[i - n]);
                    ay1 += C(n) * (f1[i + (n - 1) * sy] - f1[i - n * sy]);
                    az1 += C(n) * (f1[i + (n - 1) * sz] - f1[i - n * sz]);
                    ax2 += C(n) * (f2[i + (n - 1)] - f2[i - n]);
                    ay2 += C(n) * (f2[i + (n - 1) * sy] - f2[i - n * sy]);
                    az2 += C(n) * (f2[i + (n - 1) * sz] - f2[i - n * sz]);
                }
                const real bb = b[i];
                d1x[i] = ax1 * invdx * bb;
                d1y[i] = ay1 * invdy * bb;
                d1z[i] = az1 * invdz * bb;
                d2x[i] = ax2 * invdx * bb;
                d2y[i] = ay2 * invdy * bb;
                d2z[i] = az2 * invdz * bb;
            }
}

/* TTI combination at one point (kernels/kernel_type2_fwd.cu:196-235 calculate_result_interior): the F_aa terms
 * arrive already multiplied by rho; the mixed F_ab terms are not. */
static inline void
tti_combine(real m1, real m2, real m3, real m4, real rx, real ry, real rz, real f1xx, real f1yy, real f1zz, real f1xy,
            real f1yz, real f1zx, real f2xx, real f2yy, real f2zz, real f2xy, real f2yz, real f2zx, real* rhs3,
            real* rhs4)
{
    const real rx2 = rx * rx, ry2 = ry * ry, rz2 = rz * rz;
    const real rxy = 2 * rx * ry, ryz = 2 * ry * rz, rzx = 2 * rz * rx;
    const real v1 = rx2 * f1xx + ry2 * f1yy + rz2 * f1zz + rxy * f1xy + rzx * f1zx + ryz * f1yz;
    const real h1 = f1xx + f1yy + f1zz - v1;
    const real v2 = rx2 * f2xx + ry2 * f2yy + rz2 * f2zz + rxy * f2xy + rzx * f2zx + ryz * f2yz;
    const real h2 = f2xx + f2yy + f2zz - v2;
    *rhs3 = m1 * h1 + m4 * v1 + m2 * v2;
    *rhs4 = m2 * h1 + m3 * v2 + m4 * h2;
}

/* Rotated mixed derivative  X_ab g = inv4 * sum_n e_n [ g(+n,+n) - g(+n,-n) + g(-n,-n) - g(-n,+n) ]
 * (kernels/kernel_type2_fwd.cu:121-141 update_z_mixed, :143-164 compute_dxy) with inv4 = 1/(4 da db)
 * (kernels/kernel_utils.h:94-96). sa, sb are the strides of axes a and b. */
static inline real
mixed(const real* g, long i, long sa, long sb, real inv4)
{
    real acc = 0;
    for (int n = 1; n <= ORAD; ++n)
        acc += E(n) * ((g[i + n * sa + n * sb] - g[i + n * sa - n * sb]) + (g[i - n * sa - n * sb] - g[i - n * sa + n * sb]));
    return acc * inv4;
}

/* D+_a g(i) = invd * sum_n c_n [ g(i+n) - g(i-n+1) ]   (value at i+1/2) */
static inline real
dplus(const real* g, long i, long s, real invd)
{
    real acc = 0;
    for (int n = 1; n <= ORAD; ++n)
        acc += C(n) * (g[i + n * s] - g[i - (n - 1) * s]);
    return acc * invd;
}

/* D-_a g(i) = invd * sum_n c_n [ g(i+n-1) - g(i-n) ]   (value at i-1/2) */
static inline real
dminus(const real* g, long i, long s, real invd)
{
    real acc = 0;
    for (int n = 1; n <= ORAD; ++n)
        acc += C(n) * (g[i + (n - 1) * s] - g[i - n * s]);
    return acc * invd;
}

/* ---------------------------------------------------------------------------------------------------------------
 * a3: forward update  u += dt*rhs3, w += dt*rhs4 with F_aa = rho * D+_a (d_a f), F_ab = X_ab f.
 *     kernels/kernel_type2_fwd.cu:94-119 (update_d2z), :143-164 (dxy), :166-194 (d2x d2y), :196-235 (combine),
 *     :407-466 (z march and atomicAdd), launcher :578-665 (invdxdy = 1/(4 dx dy) etc.).
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_fwd_update(real* u, real* w, const real* d1x, const real* d2x, const real* d1y, const real* d2y, const real* d1z,
               const real* d2z, const real* f1, const real* f2, const real* irho, const real* m1, const real* m2,
               const real* m3, const real* m4, const real* rx, const real* ry, const real* rz, int ixbeg, int ixend,
               int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy, real dt, real invdx, real invdy,
               real invdz, real invdxdy, real invdydz, real invdzdx)
{
    const long sy = ldx, sz = ldx * ldy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                const real rho = 1 / irho[i];
                const real f1xx = rho * dplus(d1x, i, 1, invdx), f2xx = rho * dplus(d2x, i, 1, invdx);
                const real f1yy = rho * dplus(d1y, i, sy, invdy), f2yy = rho * dplus(d2y, i, sy, invdy);
                const real f1zz = rho * dplus(d1z, i, sz, invdz), f2zz = rho * dplus(d2z, i, sz, invdz);
                const real f1xy = mixed(f1, i, 1, sy, invdxdy), f2xy = mixed(f2, i, 1, sy, invdxdy);
                const real f1yz = mixed(f1, i, sy, sz, invdydz), f2yz = mixed(f2, i, sy, sz, invdydz);
                const real f1zx = mixed(f1, i, sz, 1, invdzdx), f2zx = mixed(f2, i, sz, 1, invdzdx);
                real r3, r4;
                tti_combine(m1[i], m2[i], m3[i], m4[i], rx[i], ry[i], rz[i], f1xx, f1yy, f1zz, f1xy, f1yz, f1zx, f2xx,
                            f2yy, f2zz, f2xy, f2yz, f2zx, &r3, &r4);
                u[i] += r3 * dt;
                w[i] += r4 * dt;
            }
}

/* ---------------------------------------------------------------------------------------------------------------
 * a4: time update with sponge  f <- A*f + B*dt*g  with A = min(ABx[ix].a, ABy[iy].a, ABz[iz].a), B likewise.
 *     kernels/kernel_base.cu:82-104 (kernel), :107-140 (interior range). AB arrays are indexed by array index.
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_loop2(real* f1, real* f2, const real* g1, const real* g2, const orc_ab* abx, const orc_ab* aby,
          const orc_ab* abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy,
          real dt)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                const real a = fminf(abx[ix].a, fminf(aby[iy].a, abz[iz].a));
                const real bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[iz].b));
                const real r1 = a * f1[i] + bb * g1[i] * dt;
                const real r2 = a * f2[i] + bb * g2[i] * dt;
                f1[i] = r1;
                f2[i] = r2;
            }
}

/* ---------------------------------------------------------------------------------------------------------------
 * Viscoacoustic memory variables (nmech mechanisms; the reference fixes nmech = 3, kernels/kernel_base.h:27).
 * M8[m], M9[m]: memory variables of sigma_H and sigma_V; M10[m]: alpha_m; wsi[m]: omega_m. Interior points only, sponge
 * A/B as in loop2.
 *   loop1 (replaces loop2):  sigma <- A sigma + B dt (u - sum_m omega_m M_m)       (kernels/kernel_base.cu:189-219)
 *   loop3:                   M_m  <- M_m + dt (alpha_m u - omega_m M_m)            (kernels/kernel_base.cu:289-325)
 *   adj_loop1:               sigma <- A sigma + B dt (u + sum_m alpha_m M_m)       (kernels/kernel_base.cu:379-405)
 *   adj_loop3:               M_m  <- M_m + dt omega_m (-u - M_m)                   (kernels/kernel_base.cu:436-460)
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_loop1(real* f1, real* f2, const real* g1, const real* g2, real* const* m8, real* const* m9, const double* wsi,
          int nmech, const orc_ab* abx, const orc_ab* aby, const orc_ab* abz, int ixbeg, int ixend, int iybeg,
          int iyend, int izbeg, int izend, long ldx, long ldy, real dt)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                const real a = fminf(abx[ix].a, fminf(aby[iy].a, abz[iz].a));
                const real bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[iz].b));
                real rx = 0, rz = 0;
                for (int m = 0; m < nmech; ++m) {
                    rx += (real)wsi[m] * m8[m][i];
                    rz += (real)wsi[m] * m9[m][i];
                }
                const real r1 = a * f1[i] + bb * dt * (g1[i] - rx);
                const real r2 = a * f2[i] + bb * dt * (g2[i] - rz);
                f1[i] = r1;
                f2[i] = r2;
            }
}

void
orc_loop3(const real* g1, const real* g2, real* const* m8, real* const* m9, real* const* m10, const double* wsi,
          int nmech, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy, real dt)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                for (int m = 0; m < nmech; ++m) {
                    const real om = (real)wsi[m], al = m10[m][i];
                    const real mx = m8[m][i], mz = m9[m][i];
                    m8[m][i] = mx + dt * (al * g1[i] - om * mx);
                    m9[m][i] = mz + dt * (al * g2[i] - om * mz);
                }
            }
}

void
orc_adj_loop1(real* f1, real* f2, const real* g1, const real* g2, real* const* m8, real* const* m9,
              real* const* m10, int nmech, const orc_ab* abx, const orc_ab* aby, const orc_ab* abz, int ixbeg,
              int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy, real dt)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                real rx = 0, rz = 0;
                for (int m = 0; m < nmech; ++m) {
                    rx += m10[m][i] * m8[m][i];
                    rz += m10[m][i] * m9[m][i];
                }
                const real a = fminf(abx[ix].a, fminf(aby[iy].a, abz[iz].a));
                const real bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[iz].b));
                const real r1 = a * f1[i] + bb * dt * (g1[i] + rx);
                const real r2 = a * f2[i] + bb * dt * (g2[i] + rz);
                f1[i] = r1;
                f2[i] = r2;
            }
}

void
orc_adj_loop3(const real* g1, const real* g2, real* const* m8, real* const* m9, const double* wsi, int nmech,
              int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy, real dt)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                for (int m = 0; m < nmech; ++m) {
                    const real om = (real)wsi[m];
                    const real mx = m8[m][i], mz = m9[m][i];
                    m8[m][i] = mx + dt * om * (-g1[i] - mx);
                    m9[m][i] = mz + dt * om * (-g2[i] - mz);
                }
            }
}

/* ---------------------------------------------------------------------------------------------------------------
 * a5: free surface mirrors at plane izbeg.
 *   sigma: plane izbeg set to 0, planes iz < izbeg get -f(2*izbeg - iz); whole xy planes (kernels/kernel_base.cu:142-187)
 *   derivatives: x, y derivatives odd about izbeg (0 at izbeg); z derivatives even, f(iz) = f(2*izbeg+1-iz) for all
 *   iz <= izbeg (kernels/kernel_type2_small.cu:5-40, host kernels/kernel_type2.cu:279-293).
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_loop4(real* f1, real* f2, long ldx, long ldy, int izbeg)
{
    const long plane = ldx * ldy;
    for (int iz = 0; iz <= izbeg; ++iz)
        for (long p = 0; p < plane; ++p) {
            const long i = (long)iz * plane + p;
            if (iz == izbeg) {
                f1[i] = 0;
                f2[i] = 0;
            } else {
                const long m = (long)(2 * izbeg - iz) * plane + p;
                f1[i] = -f1[m];
                f2[i] = -f2[m];
            }
        }
}

void
orc_loop4_derivatives(real* d1x, real* d1y, real* d1z, real* d2x, real* d2y, real* d2z, long ldx, long ldy, int izbeg)
{
    const long plane = ldx * ldy;
    for (int iz = 0; iz <= izbeg; ++iz)
        for (long p = 0; p < plane; ++p) {
            const long i = (long)iz * plane + p;
            if (iz == izbeg) {
                d1x[i] = d1y[i] = d2x[i] = d2y[i] = 0;
            } else {
                const long m = (long)(2 * izbeg - iz) * plane + p;
                d1x[i] = -d1x[m];
                d1y[i] = -d1y[m];
                d2x[i] = -d2x[m];
                d2y[i] = -d2y[m];
            }
            const long mz = (long)(2 * izbeg + 1 - iz) * plane + p;
            d1z[i] = d1z[mz];
            d2z[i] = d2z[mz];
        }
}

/* Adjoint pointwise temporaries (kernels/kernel_type2_grad_snap.cu:66-135 compute_temp_arrays;
 * kernels/kernel_type2_grad.cu:188-208 adj()):
 *   lap1 = m1 f1 + m2 f2,  hz1 = (m4 - m1) f1 - m2 f2,  lap2 = m4 f2,  hz2 = m2 f1 + (m3 - m4) f2 */
static inline void
adj_lap_hz(real f1, real f2, real m1, real m2, real m3, real m4, real* lap1, real* hz1, real* lap2, real* hz2)
{
    *lap1 = m1 * f1 + m2 * f2;
    *hz1 = (-m1 + m4) * f1 - m2 * f2;
    *lap2 = m4 * f2;
    *hz2 = m2 * f1 + (-m4 + m3) * f2;
}

/* ---------------------------------------------------------------------------------------------------------------
 * a6: adjoint derivatives. With S_{f,a} = rho*(lap_f + r_a^2 hz_f) and M_{f,ab} = 2 r_a r_b hz_f evaluated at
 *     every point (ghosts included):
 *       d_{f,a} = b * D-_a S_{f,a}            (interior)
 *       rhs_f   = sum_{a<b} X_ab M_{f,ab}      (interior)
 *     kernels/kernel_type2_grad_snap.cu:66-135 (temps), :137-160 (z derivative), :162-182 (z mixed),
 *     :187-240 (x, y, xy), :242-427 (kernel), launcher :429-477. tmp is caller scratch of 12 volumes' size
 *     (S1x S1y S1z S2x S2y S2z M1xy M1yz M1zx M2xy M2yz M2zx), each of total size ldx*ldy*ldz.
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_adj_drv(real* d1x, real* d2x, real* d1y, real* d2y, real* d1z, real* d2z, real* rhs3, real* rhs4,
            const real* f1, const real* f2, const real* b, const real* m1, const real* m2, const real* m3,
            const real* m4, const real* m5, const real* m6, const real* m7, int ixbeg, int ixend, int iybeg,
            int iyend, int izbeg, int izend, long ldx, long ldy, long ldz, real invdx, real invdy, real invdz,
            real invdxy, real invdyz, real invdzx, real* tmp)
{
    const long sy = ldx, sz = ldx * ldy, ntot = ldx * ldy * ldz;
    real* S1x = tmp + 0 * ntot;
    real* S1y = tmp + 1 * ntot;
    real* S1z = tmp + 2 * ntot;
    real* S2x = tmp + 3 * ntot;
    real* S2y = tmp + 4 * ntot;
    real* S2z = tmp + 5 * ntot;
    real* M1xy = tmp + 6 * ntot;
    real* M1yz = tmp + 7 * ntot;
    real* M1zx = tmp + 8 * ntot;
    real* M2xy = tmp + 9 * ntot;
    real* M2yz = tmp + 10 * ntot;
    real* M2zx = tmp + 11 * ntot;
    /* temporaries over the interior extended by 4 in every direction (what the reference tiles load) */
    const int r = ORAD;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg - r; iz <= izend + r; ++iz)
        for (int iy = iybeg - r; iy <= iyend + r; ++iy)
            for (int ix = ixbeg - r; ix <= ixend + r; ++ix) {
                const long i = IDX(ix, iy, iz);
                real lap1, hz1, lap2, hz2;
                adj_lap_hz(f1[i], f2[i], m1[i], m2[i], m3[i], m4[i], &lap1, &hz1, &lap2, &hz2);
                const real rho = 1 / b[i];
                const real rx = m5[i], ry = m6[i], rz = m7[i];
                S1x[i] = rho * (lap1 + rx * rx * hz1);
                S2x[i] = rho * (lap2 + rx * rx * hz2);
                S1y[i] = rho * (lap1 + ry * ry * hz1);
                S2y[i] = rho * (lap2 + ry * ry * hz2);
                S1z[i] = rho * (lap1 + rz * rz * hz1);
                S2z[i] = rho * (lap2 + rz * rz * hz2);
                M1xy[i] = 2 * rx * ry * hz1;
                M2xy[i] = 2 * rx * ry * hz2;
                M1yz[i] = 2 * ry * rz * hz1;
                M2yz[i] = 2 * ry * rz * hz2;
                M1zx[i] = 2 * rz * rx * hz1;
                M2zx[i] = 2 * rz * rx * hz2;
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                const real bb = b[i];
                d1x[i] = dminus(S1x, i, 1, invdx) * bb;
                d2x[i] = dminus(S2x, i, 1, invdx) * bb;
                d1y[i] = dminus(S1y, i, sy, invdy) * bb;
                d2y[i] = dminus(S2y, i, sy, invdy) * bb;
                d1z[i] = dminus(S1z, i, sz, invdz) * bb;
                d2z[i] = dminus(S2z, i, sz, invdz) * bb;
                rhs3[i] = mixed(M1xy, i, 1, sy, invdxy) + mixed(M1yz, i, sy, sz, invdyz) + mixed(M1zx, i, sz, 1, invdzx);
                rhs4[i] = mixed(M2xy, i, 1, sy, invdxy) + mixed(M2yz, i, sy, sz, invdyz) + mixed(M2zx, i, sz, 1, invdzx);
            }
}

/* ---------------------------------------------------------------------------------------------------------------
 * a7: adjoint update  u += dt*(sum_a D+_a d_{1,a} + rhs3), w likewise; with the sponge the reference writes
 *     u = A*u + B*dt*rhs instead (A, B = min over the three axes).
 *     kernels/kernel_type2_adj.cu:85-98 (dfdx_stg_right), :196-239 (process_z_plane), launcher :302-349.
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_adj_update(real* u, real* w, const real* d1x, const real* d2x, const real* d1y, const real* d2y, const real* d1z,
               const real* d2z, const real* rhs3, const real* rhs4, const orc_ab* abx, const orc_ab* aby,
               const orc_ab* abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy,
               real dt, real invdx, real invdy, real invdz, int sponge_active)
{
    const long sy = ldx, sz = ldx * ldy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                const real r3 = dplus(d1x, i, 1, invdx) + dplus(d1y, i, sy, invdy) + dplus(d1z, i, sz, invdz) + rhs3[i];
                const real r4 = dplus(d2x, i, 1, invdx) + dplus(d2y, i, sy, invdy) + dplus(d2z, i, sz, invdz) + rhs4[i];
                if (sponge_active) {
                    const real a = fminf(fminf(abx[ix].a, aby[iy].a), abz[iz].a);
                    const real bb = fminf(fminf(abx[ix].b, aby[iy].b), abz[iz].b) * dt;
                    u[i] = a * u[i] + bb * r3;
                    w[i] = a * w[i] + bb * r4;
                } else {
                    u[i] += r3 * dt;
                    w[i] += r4 * dt;
                }
            }
}

/* ---------------------------------------------------------------------------------------------------------------
 * a8/a13: zero-lag imaging condition  g(ix-ixbeg, ...) += adj3(ix,...)*snap3(ix-ixbeg,...) + adj4*snap4
 *     kernels/kernel_type2_small.cu:42-59 (host kernels/kernel_type2.cu:306-330); kernels/kernel_type1_gpu_grad.cu:5-22.
 *     snap and grad share the no-halo layout (sldx, sldy); the adjoint fields use (ldx, ldy).
 * finalize: g *= -dt_xcorr*2/vp with vp = sqrt(model3)   (kernels/kernel_type2_small.cu:61-76)
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_update_vp_gradient(const real* snap3, const real* snap4, real* grad, const real* adj3, const real* adj4,
                       int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long sldx, long sldy,
                       long ldx, long ldy)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long s = ((long)(iz - izbeg) * sldy + (iy - iybeg)) * sldx + (ix - ixbeg);
                const long i = IDX(ix, iy, iz);
                grad[s] += adj3[i] * snap3[s] + adj4[i] * snap4[s];
            }
}

void
orc_finalize_vp_gradient(const real* model3, real* grad, int ixbeg, int ixend, int iybeg, int iyend, int izbeg,
                         int izend, long sldx, long sldy, long ldx, long ldy, real dt_xcorr)
{
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long s = ((long)(iz - izbeg) * sldy + (iy - iybeg)) * sldx + (ix - ixbeg);
                const real vp = sqrt(model3[IDX(ix, iy, iz)]);
                grad[s] *= -dt_xcorr * 2.0 / vp;
            }
}

/* ---------------------------------------------------------------------------------------------------------------
 * a10: density gradient
 *   g += -b * sum_{f,a} [ S_{f,a} * D+_a Dfwd_{f,a} + (D-_a S_{f,a}) * Dfwd_{f,a} ]
 *   S from the adjoint sigma (as in a6), Dfwd the forward derivatives of the snapshot sigma (a9).
 *   kernels/kernel_type2_grad.cu:188-208 (adj), :210-260 (stg), :511-560 (simple kernel), launcher :614-672.
 *   The gradient pointer is NOT shifted by the interior offset: g index = (iz-izbeg)*gsz + (iy-iybeg)*gsy + (ix-ixbeg)
 *   with gsy = gldx, gsz = gldx*gldy. tmp: 6 volumes of ldx*ldy*ldz (S1x S1y S1z S2x S2y S2z).
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_rho_gradient(real* grad, const real* fd1x, const real* fd2x, const real* fd1y, const real* fd2y,
                 const real* fd1z, const real* fd2z, const real* f1, const real* f2, const real* b, const real* m1,
                 const real* m2, const real* m3, const real* m4, const real* rx, const real* ry, const real* rz,
                 int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy, long ldz,
                 long gldx, long gldy, real hx, real hy, real hz, real* tmp)
{
    const long sy = ldx, sz = ldx * ldy, ntot = ldx * ldy * ldz;
    real* S1x = tmp + 0 * ntot;
    real* S1y = tmp + 1 * ntot;
    real* S1z = tmp + 2 * ntot;
    real* S2x = tmp + 3 * ntot;
    real* S2y = tmp + 4 * ntot;
    real* S2z = tmp + 5 * ntot;
    const int r = ORAD;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg - r; iz <= izend + r; ++iz)
        for (int iy = iybeg - r; iy <= iyend + r; ++iy)
            for (int ix = ixbeg - r; ix <= ixend + r; ++ix) {
                const long i = IDX(ix, iy, iz);
                real lap1, hz1, lap2, hz2;
                adj_lap_hz(f1[i], f2[i], m1[i], m2[i], m3[i], m4[i], &lap1, &hz1, &lap2, &hz2);
                const real rho = 1 / b[i];
                S1x[i] = rho * (lap1 + rx[i] * rx[i] * hz1);
                S2x[i] = rho * (lap2 + rx[i] * rx[i] * hz2);
                S1y[i] = rho * (lap1 + ry[i] * ry[i] * hz1);
                S2y[i] = rho * (lap2 + ry[i] * ry[i] * hz2);
                S1z[i] = rho * (lap1 + rz[i] * rz[i] * hz1);
                S2z[i] = rho * (lap2 + rz[i] * rz[i] * hz2);
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                const real rhs = S1x[i] * dplus(fd1x, i, 1, hx) + S2x[i] * dplus(fd2x, i, 1, hx) +
                                 S1y[i] * dplus(fd1y, i, sy, hy) + S2y[i] * dplus(fd2y, i, sy, hy) +
                                 S1z[i] * dplus(fd1z, i, sz, hz) + S2z[i] * dplus(fd2z, i, sz, hz) +
                                 dminus(S1x, i, 1, hx) * fd1x[i] + dminus(S2x, i, 1, hx) * fd2x[i] +
                                 dminus(S1y, i, sy, hy) * fd1y[i] + dminus(S2y, i, sy, hy) * fd2y[i] +
                                 dminus(S1z, i, sz, hz) * fd1z[i] + dminus(S2z, i, sz, hz) * fd2z[i];
                const long g = ((long)(iz - izbeg) * gldy + (iy - iybeg)) * gldx + (ix - ixbeg);
                grad[g] += -b[i] * rhs;
            }
}

/* Central second derivative  invdaa * (e0 g + sum_n e_n (g(i-n) + g(i+n)))  (kernels/helper_kernels_gpu.h:120-148) */
static inline real
central2(const real* g, long i, long s, real invdaa)
{
    real acc = E(0) * g[i];
    for (int n = 1; n <= ORAD; ++n)
        acc += E(n) * (g[i - n * s] + g[i + n * s]);
    return acc * invdaa;
}

/* ---------------------------------------------------------------------------------------------------------------
 * a11: constant-density TTI forward (type1). F_aa = central2_a f, F_ab = X_ab f, same combination as a3 without rho.
 *   mode 0: u += dt*rhs; mode 1: u = A*u + B*dt*rhs (sponge); mode 2 (calc_snap): out(no halo) = rhs.
 *   kernels/kernel_type1_gpu_fwd.cu:9-71 (simple), :225-437 (optimised; its x-z mixed stencil is the one used here,
 *   see the quirk note on drv31 in the design notes), launcher :442-545.
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_t1_fwd(real* u, real* w, const real* f1, const real* f2, const real* m1, const real* m2, const real* m3,
           const real* m4, const real* m5, const real* m6, const real* m7, const orc_ab* abx, const orc_ab* aby,
           const orc_ab* abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy,
           long oldx, long oldy, real dt, real invdxx, real invdyy, real invdzz, real invdxy, real invdyz,
           real invdzx, int mode)
{
    const long sy = ldx, sz = ldx * ldy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                real r3, r4;
                tti_combine(m1[i], m2[i], m3[i], m4[i], m5[i], m6[i], m7[i], central2(f1, i, 1, invdxx),
                            central2(f1, i, sy, invdyy), central2(f1, i, sz, invdzz), mixed(f1, i, 1, sy, invdxy),
                            mixed(f1, i, sy, sz, invdyz), mixed(f1, i, sz, 1, invdzx), central2(f2, i, 1, invdxx),
                            central2(f2, i, sy, invdyy), central2(f2, i, sz, invdzz), mixed(f2, i, 1, sy, invdxy),
                            mixed(f2, i, sy, sz, invdyz), mixed(f2, i, sz, 1, invdzx), &r3, &r4);
                if (mode == 2) {
                    const long o = ((long)(iz - izbeg) * oldy + (iy - iybeg)) * oldx + (ix - ixbeg);
                    u[o] = r3;
                    w[o] = r4;
                } else if (mode == 1) {
                    const real a = fminf(abx[ix].a, fminf(aby[iy].a, abz[iz].a));
                    const real bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[iz].b));
                    u[i] = a * u[i] + bb * r3 * dt;
                    w[i] = a * w[i] + bb * r4 * dt;
                } else {
                    u[i] += r3 * dt;
                    w[i] += r4 * dt;
                }
            }
}

/* ---------------------------------------------------------------------------------------------------------------
 * a12: constant-density TTI adjoint (type1).
 *   u += dt*( sum_a central2_a(T_{1,a}) + sum_{a<b} X_ab(M_{1,ab}) ),  T_{f,a} = lap_f + r_a^2 hz_f (no rho),
 *   M_{f,ab} = 2 r_a r_b hz_f;  sponge: u = A*u + B*dt*rhs.
 *   kernels/kernel_type1_gpu_adj.cu:9-51 (simple), :235-378 (optimised), kernels/helper_kernels_gpu.h:186-418.
 *   tmp: 12 volumes as in orc_adj_drv (T1x T1y T1z T2x T2y T2z M1xy M1yz M1zx M2xy M2yz M2zx).
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_t1_adj(real* u, real* w, const real* f1, const real* f2, const real* m1, const real* m2, const real* m3,
           const real* m4, const real* m5, const real* m6, const real* m7, const orc_ab* abx, const orc_ab* aby,
           const orc_ab* abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, long ldx, long ldy,
           long ldz, real dt, real invdxx, real invdyy, real invdzz, real invdxy, real invdyz, real invdzx,
           int use_sponge, real* tmp)
{
    const long sy = ldx, sz = ldx * ldy, ntot = ldx * ldy * ldz;
    real* T[12];
    for (int k = 0; k < 12; ++k)
        T[k] = tmp + k * ntot;
    const int r = ORAD;
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg - r; iz <= izend + r; ++iz)
        for (int iy = iybeg - r; iy <= iyend + r; ++iy)
            for (int ix = ixbeg - r; ix <= ixend + r; ++ix) {
                const long i = IDX(ix, iy, iz);
                real lap1, hz1, lap2, hz2;
                adj_lap_hz(f1[i], f2[i], m1[i], m2[i], m3[i], m4[i], &lap1, &hz1, &lap2, &hz2);
                const real rx = m5[i], ry = m6[i], rz = m7[i];
                T[0][i] = lap1 + rx * rx * hz1;
                T[1][i] = lap1 + ry * ry * hz1;
                T[2][i] = lap1 + rz * rz * hz1;
                T[3][i] = lap2 + rx * rx * hz2;
                T[4][i] = lap2 + ry * ry * hz2;
                T[5][i] = lap2 + rz * rz * hz2;
                T[6][i] = 2 * rx * ry * hz1;
                T[7][i] = 2 * ry * rz * hz1;
                T[8][i] = 2 * rz * rx * hz1;
                T[9][i] = 2 * rx * ry * hz2;
                T[10][i] = 2 * ry * rz * hz2;
                T[11][i] = 2 * rz * rx * hz2;
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int iz = izbeg; iz <= izend; ++iz)
        for (int iy = iybeg; iy <= iyend; ++iy)
            for (int ix = ixbeg; ix <= ixend; ++ix) {
                const long i = IDX(ix, iy, iz);
                const real r3 = central2(T[0], i, 1, invdxx) + central2(T[1], i, sy, invdyy) +
                                central2(T[2], i, sz, invdzz) + mixed(T[6], i, 1, sy, invdxy) +
                                mixed(T[7], i, sy, sz, invdyz) + mixed(T[8], i, sz, 1, invdzx);
                const real r4 = central2(T[3], i, 1, invdxx) + central2(T[4], i, sy, invdyy) +
                                central2(T[5], i, sz, invdzz) + mixed(T[9], i, 1, sy, invdxy) +
                                mixed(T[10], i, sy, sz, invdyz) + mixed(T[11], i, sz, 1, invdzx);
                if (use_sponge) {
                    const real a = fminf(abx[ix].a, fminf(aby[iy].a, abz[iz].a));
                    const real bb = fminf(abx[ix].b, fminf(aby[iy].b, abz[iz].b));
                    u[i] = a * u[i] + bb * r3 * dt;
                    w[i] = a * w[i] + bb * r4 * dt;
                } else {
                    u[i] += r3 * dt;
                    w[i] += r4 * dt;
                }
            }
}

/* ---------------------------------------------------------------------------------------------------------------
 * a14: model build  m1 = vp^2(1+2eps), m2 = sqrt((vp^2-vs^2)(vp^2(1+2delta)-vs^2)), m3 = vp^2, m4 = vs^2,
 *      m5..m7 = (sin t cos p, sin t sin p, cos t), b = 1/rho. Per-point inputs (the reference uses constants).
 *      simulator2/simulator2SetUp.cu:10-29. Pressure P = -(2 f1 + f2)/3 (simulator2/simulator2SetUp.cu:31-42).
 * ------------------------------------------------------------------------------------------------------------- */
void
orc_init_cijs(long n, const real* vp, const real* eps, const real* delta, const real* vs, const real* theta,
              const real* phi, const real* rho, real* m1, real* m2, real* m3, real* m4, real* m5, real* m6, real* m7,
              real* b)
{
    for (long i = 0; i < n; ++i) {
        const real v = vp[i], s = vs[i];
        m1[i] = v * v * (1 + 2 * eps[i]);
        m2[i] = sqrt((v * v - s * s) * (v * v * (1 + 2 * delta[i]) - s * s));
        m3[i] = v * v;
        m4[i] = s * s;
        m5[i] = sin(theta[i]) * cos(phi[i]);
        m6[i] = sin(theta[i]) * sin(phi[i]);
        m7[i] = cos(theta[i]);
        b[i] = 1 / rho[i];
    }
}

void
orc_pressure(long n, const real* f1, const real* f2, real* p)
{
    for (long i = 0; i < n; ++i)
        p[i] = -(2 * f1[i] + f2[i]) / 3;
}
