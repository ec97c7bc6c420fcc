// Launchers of the order-16 (and order-12 density-gradient) operators in order_kernels.hip, called by the C entry
// points of type2_kernels.hip and type1_kernels.hip when their `order` argument is not 8. Arguments mean what the C
// entry points' do; each returns a TTI_* status.
#pragma once

namespace tti {

int order_fwd_drv(int order, const float* f1, const float* f2, const float* b, float* d1x, float* d1y, float* d1z,
                  float* d2x, float* d2y, float* d2z, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                  float invdx, float invdy, float invdz, int ldimx, int ldimy, void* stream);
int order_fwd_update(int order, float* u, float* w, const float* d1x, const float* d2x, const float* d1y,
                     const float* d2y, const float* d1z, const float* d2z, const float* f1, const float* f2,
                     const float* irho, const float* m1, const float* m2, const float* m3, const float* m4,
                     const float* rx, const float* ry, const float* rz, int ixbeg, int ixend, int iybeg, int iyend,
                     int izbeg, int izend, int ldimx, int ldimy, float dt, double dx, double dy, double dz,
                     void* stream);
int order_adj_drv(int order, float* d1x, float* d2x, float* d1y, float* d2y, float* d1z, float* d2z, float* r3,
                  float* r4, const float* f1, const float* f2, const float* b, const float* m1, const float* m2,
                  const float* m3, const float* m4, const float* rx, const float* ry, const float* rz, int ixbeg,
                  int ixend, int iybeg, int iyend, int izbeg, int izend, float invdxy, float invdyz, float invdzx,
                  int ldimx, int ldimy, float invdx, float invdy, float invdz, void* stream);
int order_adj_update(int order, float* u, float* w, const float* d1x, const float* d2x, const float* d1y,
                     const float* d2y, const float* d1z, const float* d2z, const float* r3, const float* r4,
                     const void* abx, const void* aby, const void* abz, int ixbeg, int ixend, int iybeg, int iyend,
                     int izbeg, int izend, int ldimx, int ldimy, float dt, double dx, double dy, double dz,
                     int sponge_active, void* stream);
int order_rho_grad(int order, float* grad, const float* e1x, const float* e2x, const float* e1y, const float* e2y,
                   const float* e1z, const float* e2z, const float* f1, const float* f2, const float* irho,
                   const float* m1, const float* m2, const float* m3, const float* m4, const float* rx, const float* ry,
                   const float* rz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx,
                   int ldimy, int grad_ldimx, int grad_ldimy, double dx, double dy, double dz, void* stream);
int order_t1_fwd(int order, float* u, float* w, const float* f1, const float* f2, const float* m1, const float* m2,
                 const float* m3, const float* m4, const float* m5, const float* m6, const float* m7, const void* abx,
                 const void* aby, const void* abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                 int ldimx_in, int ldimy_in, int ldimx_out, int ldimy_out, float dt, double dx, double dy, double dz,
                 int mode, void* stream);
int order_t1_adj(int order, float* u, float* w, const float* f1, const float* f2, const float* m1, const float* m2,
                 const float* m3, const float* m4, const float* m5, const float* m6, const float* m7, const void* abx,
                 const void* aby, const void* abz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend,
                 int ldimx, int ldimy, float dt, double dx, double dy, double dz, int use_sponge, void* stream);

} // namespace tti
