/*
 * TEST INFRASTRUCTURE ONLY -- time-step compositions of the oracle primitives (see tti_oracle.c for the rules on
 * who may load this library).
 *
 * One forward step, as simulator2SetUp::execute orders it (simulator2/simulator2SetUp.cu:530-649):
 *   sigma <- A sigma + B dt u            (kernel_loop2)
 *   [free-surface mirror of sigma]       (kernel_loop4)
 *   d = b D- sigma                       (fwd_main_loop)
 *   [free-surface mirror of d]           (kernel_loop4_derivative)
 *   u += dt L(sigma, d)                  (fwd_main_loop_2)
 * One adjoint step (simulator2/simulator2SetUp.cu:662-737):
 *   (d, R) = adjDerivs(sigma); [mirror d]; u += dt (D+ d + R) (sponge form when active);
 *   sigma <- A sigma + B dt u; [mirror sigma]
 * The single-volume case has no halo exchange; ghosts keep whatever the caller put there (zeros).
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
typedef struct
{
    float a, b;
} orc_ab;

/* primitives from tti_oracle.c */
void orc_fwd_drv(const real*, const real*, const real*, real*, real*, real*, real*, real*, real*, int, int, int, int,
                 int, int, real, real, real, long, long);
void orc_fwd_update(real*, real*, const real*, const real*, const real*, const real*, const real*, const real*,
                    const real*, const real*, const real*, const real*, const real*, const real*, const real*,
                    const real*, const real*, const real*, int, int, int, int, int, int, long, long, real, real, real,
                    real, real, real, real);
void orc_loop2(real*, real*, const real*, const real*, const orc_ab*, const orc_ab*, const orc_ab*, int, int, int, int,
               int, int, long, long, real);
void orc_loop4(real*, real*, long, long, int);
void orc_loop4_derivatives(real*, real*, real*, real*, real*, real*, long, long, int);
void orc_adj_drv(real*, real*, real*, real*, real*, real*, real*, real*, const real*, const real*, const real*,
                 const real*, const real*, const real*, const real*, const real*, const real*, const real*, int, int,
                 int, int, int, int, long, long, long, real, real, real, real, real, real, real*);
void orc_adj_update(real*, real*, const real*, const real*, const real*, const real*, const real*, const real*,
                    const real*, const real*, const orc_ab*, const orc_ab*, const orc_ab*, int, int, int, int, int, int,
                    long, long, real, real, real, real, int);

/* Geometry and constants shared by the step functions. */
typedef struct
{
    int ixbeg, ixend, iybeg, iyend, izbeg, izend;
    long ldx, ldy, ldz;
    real dt;
    double dx, dy, dz;
    int free_surface;
    int sponge_active;
} orc_geom;

static void
fd_consts(const orc_geom* g, real* invdx, real* invdy, real* invdz, real* i4xy, real* i4yz, real* i4zx)
{
    /* kernel_utils::compute_fd_const stores these as float (kernels/kernel_utils.h:83-97) */
    *invdx = (float)(1.0 / g->dx);
    *invdy = (float)(1.0 / g->dy);
    *invdz = (float)(1.0 / g->dz);
    *i4xy = (float)(1.0 / (4 * g->dx * g->dy));
    *i4yz = (float)(1.0 / (4 * g->dy * g->dz));
    *i4zx = (float)(1.0 / (4 * g->dx * g->dz));
}

/* models: m[0..7] = m1 m2 m3 m4 rx ry rz b ; d[0..5] = d1x d1y d1z d2x d2y d2z (scratch, ghosts zero) */
void
orc_step_fwd(const orc_geom* g, real* f1, real* f2, real* u, real* w, const real* const* m, real* const* d,
             const orc_ab* abx, const orc_ab* aby, const orc_ab* abz)
{
    real ix, iy, iz, a, b, c;
    fd_consts(g, &ix, &iy, &iz, &a, &b, &c);
    orc_loop2(f1, f2, u, w, abx, aby, abz, g->ixbeg, g->ixend, g->iybeg, g->iyend, g->izbeg, g->izend, g->ldx, g->ldy,
              g->dt);
    if (g->free_surface)
        orc_loop4(f1, f2, g->ldx, g->ldy, g->izbeg);
    orc_fwd_drv(f1, f2, m[7], d[0], d[1], d[2], d[3], d[4], d[5], g->ixbeg, g->ixend, g->iybeg, g->iyend, g->izbeg,
                g->izend, ix, iy, iz, g->ldx, g->ldy);
    if (g->free_surface)
        orc_loop4_derivatives(d[0], d[1], d[2], d[3], d[4], d[5], g->ldx, g->ldy, g->izbeg);
    orc_fwd_update(u, w, d[0], d[3], d[1], d[4], d[2], d[5], f1, f2, m[7], m[0], m[1], m[2], m[3], m[4], m[5], m[6],
                   g->ixbeg, g->ixend, g->iybeg, g->iyend, g->izbeg, g->izend, g->ldx, g->ldy, g->dt, ix, iy, iz, a, b,
                   c);
}

/* d[0..7] = d1x d1y d1z d2x d2y d2z rhs3 rhs4 ; tmp = 12 volumes of scratch */
void
orc_step_adj(const orc_geom* g, real* f1, real* f2, real* u, real* w, const real* const* m, real* const* d,
             const orc_ab* abx, const orc_ab* aby, const orc_ab* abz, real* tmp)
{
    real ix, iy, iz, a, b, c;
    fd_consts(g, &ix, &iy, &iz, &a, &b, &c);
    orc_adj_drv(d[0], d[3], d[1], d[4], d[2], d[5], d[6], d[7], f1, f2, m[7], m[0], m[1], m[2], m[3], m[4], m[5], m[6],
                g->ixbeg, g->ixend, g->iybeg, g->iyend, g->izbeg, g->izend, g->ldx, g->ldy, g->ldz, ix, iy, iz, a, b,
                c, tmp);
    if (g->free_surface)
        orc_loop4_derivatives(d[0], d[1], d[2], d[3], d[4], d[5], g->ldx, g->ldy, g->izbeg);
    orc_adj_update(u, w, d[0], d[3], d[1], d[4], d[2], d[5], d[6], d[7], abx, aby, abz, g->ixbeg, g->ixend, g->iybeg,
                   g->iyend, g->izbeg, g->izend, g->ldx, g->ldy, g->dt, ix, iy, iz, g->sponge_active);
    orc_loop2(f1, f2, u, w, abx, aby, abz, g->ixbeg, g->ixend, g->iybeg, g->iyend, g->izbeg, g->izend, g->ldx, g->ldy,
              g->dt);
    if (g->free_surface)
        orc_loop4(f1, f2, g->ldx, g->ldy, g->izbeg);
}

/*
 * CPU baseline: nt forward + nt adjoint steps of the TTI operator on an n^3 interior with 4-point ghosts, FP of this
 * build, all OpenMP threads. Synthetic smooth model (vp 1500..4500 with depth, eps 0.1, delta 0.05, tilt 30 deg,
 * azimuth 45 deg, rho 1000) and a smooth initial field. Returns elapsed seconds of the time loops only
 * (allocation and initialisation excluded) and the number of threads used in *threads.
 */
double
orc_bench_fwdadj(int n, int nt, int* threads)
{
    const int G = 4;
    const long ldx = ((n + 2 * G + 31) / 32) * 32, ldy = n + 2 * G, ldz = n + 2 * G;
    const long ntot = ldx * ldy * ldz;
    orc_geom g = { G, G + n - 1, G, G + n - 1, G, G + n - 1, ldx, ldy, ldz, 5e-4f, 10.0, 10.0, 10.0, 0, 0 };
    real* buf = (real*)calloc((size_t)ntot * (4 + 8 + 8 + 12), sizeof(real));
    if (!buf)
        return -1.0;
    real* f1 = buf;
    real* f2 = buf + ntot;
    real* u = buf + 2 * ntot;
    real* w = buf + 3 * ntot;
    real* m[8];
    real* d[8];
    for (int k = 0; k < 8; ++k) {
        m[k] = buf + (4 + k) * ntot;
        d[k] = buf + (12 + k) * ntot;
    }
    real* tmp = buf + 20 * ntot;
    orc_ab* ab = (orc_ab*)malloc(sizeof(orc_ab) * (size_t)(ldx + ldy + ldz));
    for (long i = 0; i < ldx + ldy + ldz; ++i)
        ab[i].a = ab[i].b = 1.0f;
#pragma omp parallel for
    for (long iz = 0; iz < ldz; ++iz)
        for (long iy = 0; iy < ldy; ++iy)
            for (long ix = 0; ix < ldx; ++ix) {
                const long i = (iz * ldy + iy) * ldx + ix;
                const real vp = 1500 + 3000 * (real)iz / (real)ldz;
                const real th = 0.5236, ph = 0.7854;
                m[0][i] = vp * vp * 1.2;
                m[1][i] = vp * vp * sqrt(1.1);
                m[2][i] = vp * vp;
                m[3][i] = 0;
                m[4][i] = sin(th) * cos(ph);
                m[5][i] = sin(th) * sin(ph);
                m[6][i] = cos(th);
                m[7][i] = 1e-3;
                const int inside = ix >= G && ix < G + n && iy >= G && iy < G + n && iz >= G && iz < G + n;
                if (inside) {
                    const real rx = (ix - ldx / 2.0) / n, ry = (iy - ldy / 2.0) / n, rz = (iz - ldz / 2.0) / n;
                    f1[i] = exp(-50 * (rx * rx + ry * ry + rz * rz));
                    f2[i] = 0.5 * f1[i];
                }
            }
    *threads = omp_get_max_threads();
    const double t0 = omp_get_wtime();
    for (int it = 0; it < nt; ++it)
        orc_step_fwd(&g, f1, f2, u, w, (const real* const*)m, d, ab, ab + ldx, ab + ldx + ldy);
    for (int it = 0; it < nt; ++it)
        orc_step_adj(&g, f1, f2, u, w, (const real* const*)m, d, ab, ab + ldx, ab + ldx + ldy, tmp);
    const double t1 = omp_get_wtime();
    free(ab);
    free(buf);
    return t1 - t0;
}
