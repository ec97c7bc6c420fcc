// Entry points of type2_kernels.hip for the other translation units of the library (not part of the C ABI).
#pragma once

namespace tti {

// launch_update_rho_gradient_rho (order 8) accumulating into an FP64 gradient (same layout and arguments)
int rho_gradient_f64(double* grad, const float* fwd_field1_dx, const float* fwd_field2_dx, const float* fwd_field1_dy,
                     const float* fwd_field2_dy, const float* fwd_field1_dz, const float* fwd_field2_dz,
                     const float* field1, const float* field2, const float* irho, const float* model1,
                     const float* model2, const float* model3, const float* model4, const float* rx, const float* ry,
                     const float* rz, int ixbeg, int ixend, int iybeg, int iyend, int izbeg, int izend, int ldimx,
                     int ldimy, int grad_ldimx, int grad_ldimy, double dx, double dy, double dz, void* stream);

} // namespace tti
