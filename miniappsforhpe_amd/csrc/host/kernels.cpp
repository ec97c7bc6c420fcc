// Operator classes: argument marshalling from the volumes to the C ABI of libtti_hip.so (see kernels.h).
#include "kernels.h"

#include <algorithm>

#include "tti_hip.h"

namespace {

using gpu_vols = std::vector<cart_volume_regular_gpu*>;

// Lazy cast of the caller's volume array (filled after the operator's construction), as the reference does in
// ensure_volumes (kernels/kernel_type2.cu:45-123, kernels/kernel_type1_gpu.cpp:27-90). Every volume must be
// 128-byte aligned with a 128-byte multiple row length: the kernels read rows as 16-byte vectors.
void
cast_vols(cart_volume<float>** in, int nsub, gpu_vols& out)
{
    out.clear();
    if (!in)
        return;
    out.resize(nsub, nullptr);
    std::transform(in, in + nsub, out.begin(), [](cart_volume<float>* cv) -> cart_volume_regular_gpu* {
        if (!cv)
            return nullptr;
        auto* g = cv->as<cart_volume_regular_gpu>();
        if (!g->isAligned())
            throw std::runtime_error("operator volume is not 128-byte aligned");
        return g;
    });
}

struct range
{
    int ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy;
};

range
range_of(const cart_volume_regular* v)
{
    return { v->ax1()->ibeg, v->ax1()->iend, v->ax2()->ibeg, v->ax2()->iend,
             v->ax3()->ibeg, v->ax3()->iend, v->ax1()->ntot, v->ax2()->ntot };
}

float*
data(cart_volume_regular_gpu* v)
{
    if (!v)
        throw std::runtime_error("operator volume is not set");
    return v->getData();
}

float*
gdata(cart_volume<float>* v)
{
    return data(v->as<cart_volume_regular_gpu>());
}

void
check(int rc, const char* what)
{
    tti::abi_check(rc, what);
}

void
no_pml(bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    if (pml_x || pml_y || pml_z || pml_any)
        throw std::runtime_error("PML boundaries are not supported: use the sponge (bc_opt = _BDY_SPONGE)");
}

} // namespace

// ---- sponge holders ------------------------------------------------------------------------------------------------

sponge_coef3d_gpu::~sponge_coef3d_gpu()
{
    for (auto* v : { &x_AB_, &y_AB_, &z_AB_ })
        for (float2* p : *v)
            if (p)
                (void)hipFree(p);
}

void
sponge_coef3d_gpu::setCoeffs(const sponge_coef3d& cpu)
{
    // utils/sponge_coef3d_gpu.h:37-58: set once, one {A, B} array per axis and sub-volume, active when any
    // coefficient differs from 1
    if (nsubs != 0)
        throw std::runtime_error("sponge_coef3d_gpu: coefficients already set");
    nsubs = (int)cpu.y_A_.size();
    if (nsubs <= 0 || cpu.x_A_.size() != (size_t)nsubs || cpu.z_A_.size() != (size_t)nsubs ||
        cpu.x_B_.size() != (size_t)nsubs || cpu.y_B_.size() != (size_t)nsubs || cpu.z_B_.size() != (size_t)nsubs)
        throw std::runtime_error("sponge_coef3d_gpu: inconsistent sub-volume count");
    auto load = [this](const std::vector<float>& a, const std::vector<float>& b) -> float2* {
        if (a.empty())
            return nullptr;
        if (a.size() != b.size())
            throw std::runtime_error("sponge_coef3d_gpu: A and B lengths differ");
        std::vector<float2> ab(a.size());
        for (size_t i = 0; i < a.size(); ++i) {
            active_coefs = active_coefs || a[i] != 1.0f || b[i] != 1.0f;
            ab[i] = make_float2(a[i], b[i]);
        }
        float2* d = nullptr;
        tti::hip_check(hipMalloc((void**)&d, ab.size() * sizeof(float2)), "sponge_coef3d_gpu: hipMalloc");
        tti::hip_check(hipMemcpy(d, ab.data(), ab.size() * sizeof(float2), hipMemcpyHostToDevice),
                       "sponge_coef3d_gpu: hipMemcpy");
        return d;
    };
    for (int i = 0; i < nsubs; ++i) {
        x_AB_.push_back(load(cpu.x_A_[i], cpu.x_B_[i]));
        y_AB_.push_back(load(cpu.y_A_[i], cpu.y_B_[i]));
        z_AB_.push_back(load(cpu.z_A_[i], cpu.z_B_[i]));
    }
}

std::array<float2 const*, 3>
sponge_coef3d_gpu::getSpongeCoef(int isub) const
{
    if (isub < 0 || isub >= nsubs)
        throw std::runtime_error("sponge_coef3d_gpu: sub-volume index out of range");
    return { x_AB_[isub], y_AB_[isub], z_AB_[isub] };
}

// ---- kernel_abstract_base / kernel_base ----------------------------------------------------------------------------

void
kernel_abstract_base::kernel_loop1_inner(int, hipStream_t)
{
    throw std::runtime_error("kernel_loop1: this operator has no viscoacoustic memory variables");
}
void
kernel_abstract_base::kernel_loop3_inner(int, hipStream_t)
{
    throw std::runtime_error("kernel_loop3: this operator has no viscoacoustic memory variables");
}
void
kernel_abstract_base::adj_kernel_loop1_inner(int, hipStream_t)
{
    throw std::runtime_error("adj_kernel_loop1: this operator has no viscoacoustic memory variables");
}
void
kernel_abstract_base::adj_kernel_loop3_inner(int, hipStream_t)
{
    throw std::runtime_error("adj_kernel_loop3: this operator has no viscoacoustic memory variables");
}

kernel_base::kernel_base(int sp_order, int bc_opt, int nsub_, float dt_, double, double, double, bool update_born_,
                         cart_volume<float>**& field1, cart_volume<float>**& field2, cart_volume<float>**& field3,
                         cart_volume<float>**& field4, cart_volume<float>**& field11, cart_volume<float>**& field21,
                         cart_volume<float>**& field31, cart_volume<float>**& field41,
                         const sponge_coef3d_gpu& sponge_coef_, std::vector<cart_volume<float>**>& Model8_,
                         std::vector<cart_volume<float>**>& Model9_, std::vector<cart_volume<float>**>& Model81_,
                         std::vector<cart_volume<float>**>& Model91_, std::vector<float>& wsi_,
                         std::vector<cart_volume<float>**>& Model10_)
    : kernel_abstract_base(sp_order, bc_opt), nsub(nsub_), dt(dt_), update_born(update_born_), field1_orig(field1),
      field2_orig(field2), field3_orig(field3), field4_orig(field4), field11_orig(field11), field21_orig(field21),
      field31_orig(field31), field41_orig(field41), sponge_coef(sponge_coef_), Model8_orig(Model8_),
      Model9_orig(Model9_), Model81_orig(Model81_), Model91_orig(Model91_), Model10_orig(Model10_), wsi(wsi_)
{
}

void
kernel_base::ensure_volumes()
{
    if (cast_done)
        return;
    cast_vols(field1_orig, nsub, field1);
    cast_vols(field2_orig, nsub, field2);
    cast_vols(field3_orig, nsub, field3);
    cast_vols(field4_orig, nsub, field4);
    if (update_born) {
        // kernels/kernel_base.cu:52-57
        if (!field11_orig || !field21_orig || !field31_orig || !field41_orig)
            throw std::runtime_error("kernel_base: update_born needs the Born fields field11..field41");
        cast_vols(field11_orig, nsub, field11);
        cast_vols(field21_orig, nsub, field21);
        cast_vols(field31_orig, nsub, field31);
        cast_vols(field41_orig, nsub, field41);
    }
    cast_done = true;
}

void
kernel_base::kernel_loop2_inner(int isub, hipStream_t stream)
{
    // kernels/kernel_base.cu:106-140
    const range r = range_of(field1[isub]);
    const auto ab = sponge_coef.getSpongeCoef(isub);
    check(::kernel_loop2_inner(data(field1[isub]), data(field2[isub]), data(field3[isub]), data(field4[isub]), ab[0],
                               ab[1], ab[2], r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, dt,
                               stream),
          "kernel_loop2");
    if (update_born) // kernels/kernel_base.cu:128-137: the same update of the Born fields
        check(::kernel_loop2_inner(data(field11[isub]), data(field21[isub]), data(field31[isub]), data(field41[isub]),
                                   ab[0], ab[1], ab[2], r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx,
                                   r.ldimy, dt, stream),
              "kernel_loop2 (Born fields)");
}

// Memory variables are cast on first use: the owner may fill them after construction (as the reference's driver
// does), and drivers that never call the viscoacoustic loops need not allocate them.
void
kernel_base::ensure_memory_variables()
{
    if (mech_cast_done)
        return;
    if (Model8_orig.size() < (size_t)nmech || Model9_orig.size() < (size_t)nmech ||
        Model10_orig.size() < (size_t)nmech || wsi.size() < (size_t)nmech)
        throw std::runtime_error("kernel_base: the viscoacoustic loops need nmech = 3 memory variables and omegas");
    for (int m = 0; m < nmech; ++m) {
        if (!Model8_orig[m] || !Model9_orig[m] || !Model10_orig[m])
            throw std::runtime_error("kernel_base: memory-variable volumes are not allocated");
        cast_vols(Model8_orig[m], nsub, Model8[m]);
        cast_vols(Model9_orig[m], nsub, Model9[m]);
        cast_vols(Model10_orig[m], nsub, Model10[m]);
        if (update_born) { // kernels/kernel_base.cu:73-76
            if (Model81_orig.size() < (size_t)nmech || Model91_orig.size() < (size_t)nmech || !Model81_orig[m] ||
                !Model91_orig[m])
                throw std::runtime_error("kernel_base: update_born needs the Born memory variables Model81, Model91");
            cast_vols(Model81_orig[m], nsub, Model81[m]);
            cast_vols(Model91_orig[m], nsub, Model91[m]);
        }
    }
    mech_cast_done = true;
}

kernel_base::mech_ptrs
kernel_base::mech(int isub) const
{
    mech_ptrs p;
    for (int m = 0; m < nmech; ++m) {
        p.m8[m] = data(Model8[m][isub]);
        p.m9[m] = data(Model9[m][isub]);
        p.m10[m] = data(Model10[m][isub]);
        p.w[m] = wsi[m];
    }
    return p;
}

kernel_base::mech_ptrs
kernel_base::born_mech(int isub) const
{
    mech_ptrs p = mech(isub);
    for (int m = 0; m < nmech; ++m) {
        p.m8[m] = data(Model81[m][isub]);
        p.m9[m] = data(Model91[m][isub]);
    }
    return p;
}

void
kernel_base::kernel_loop1_inner(int isub, hipStream_t stream)
{
    // kernels/kernel_base.cu:189-272
    ensure_memory_variables();
    const range r = range_of(field1[isub]);
    const auto ab = sponge_coef.getSpongeCoef(isub);
    const mech_ptrs p = mech(isub);
    const float* m8[nmech] = { p.m8[0], p.m8[1], p.m8[2] };
    const float* m9[nmech] = { p.m9[0], p.m9[1], p.m9[2] };
    check(::kernel_loop1_inner(data(field1[isub]), data(field2[isub]), data(field3[isub]), data(field4[isub]), m8, m9,
                               p.w, nmech, ab[0], ab[1], ab[2], r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend,
                               r.ldimx, r.ldimy, dt, stream),
          "kernel_loop1");
    if (update_born) { // kernels/kernel_base.cu:256-267: the Born fields with the Born memory variables
        const mech_ptrs q = born_mech(isub);
        const float* m81[nmech] = { q.m8[0], q.m8[1], q.m8[2] };
        const float* m91[nmech] = { q.m9[0], q.m9[1], q.m9[2] };
        check(::kernel_loop1_inner(data(field11[isub]), data(field21[isub]), data(field31[isub]), data(field41[isub]),
                                   m81, m91, q.w, nmech, ab[0], ab[1], ab[2], r.ixbeg, r.ixend, r.iybeg, r.iyend,
                                   r.izbeg, r.izend, r.ldimx, r.ldimy, dt, stream),
              "kernel_loop1 (Born fields)");
    }
}

void
kernel_base::kernel_loop3_inner(int isub, hipStream_t stream)
{
    // kernels/kernel_base.cu:274-364: with update_born one pass updates both sets of memory variables
    ensure_memory_variables();
    const range r = range_of(field1[isub]);
    const mech_ptrs p = mech(isub);
    if (update_born) {
        const mech_ptrs q = born_mech(isub);
        check(::kernel_loop3_inner_born(data(field3[isub]), data(field4[isub]), data(field31[isub]),
                                        data(field41[isub]), p.m8, p.m9, q.m8, q.m9, p.m10, p.w, nmech, r.ixbeg,
                                        r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, dt, stream),
              "kernel_loop3 (with the Born fields)");
        return;
    }
    check(::kernel_loop3_inner(data(field3[isub]), data(field4[isub]), p.m8, p.m9, p.m10, p.w, nmech, r.ixbeg,
                               r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, dt, stream),
          "kernel_loop3");
}

void
kernel_base::adj_kernel_loop1_inner(int isub, hipStream_t stream)
{
    // kernels/kernel_base.cu:366-431
    ensure_memory_variables();
    const range r = range_of(field1[isub]);
    const auto ab = sponge_coef.getSpongeCoef(isub);
    const mech_ptrs p = mech(isub);
    const float* m8[nmech] = { p.m8[0], p.m8[1], p.m8[2] };
    const float* m9[nmech] = { p.m9[0], p.m9[1], p.m9[2] };
    check(::adj_kernel_loop1_inner(data(field1[isub]), data(field2[isub]), data(field3[isub]), data(field4[isub]), m8,
                                   m9, p.m10, nmech, ab[0], ab[1], ab[2], r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg,
                                   r.izend, r.ldimx, r.ldimy, dt, stream),
          "adj_kernel_loop1");
}

void
kernel_base::adj_kernel_loop3_inner(int isub, hipStream_t stream)
{
    // kernels/kernel_base.cu:433-492
    ensure_memory_variables();
    const range r = range_of(field1[isub]);
    const mech_ptrs p = mech(isub);
    check(::adj_kernel_loop3_inner(data(field3[isub]), data(field4[isub]), p.m8, p.m9, p.w, nmech, r.ixbeg, r.ixend,
                                   r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, dt, stream),
          "adj_kernel_loop3");
}

void
kernel_base::kernel_loop4_inner(int isub, hipStream_t stream)
{
    // kernels/kernel_base.cu:164-187; restricted to the sub-volumes whose top face is the free surface when the
    // owner says which those are
    if (!free_surface_.empty() && !free_surface_.at(isub))
        return;
    const range r = range_of(field1[isub]);
    check(::kernel_loop4_inner(data(field1[isub]), data(field2[isub]), r.ldimx, r.ldimy, r.izbeg, stream),
          "kernel_loop4");
    if (update_born) // kernels/kernel_base.cu:182-186
        check(::kernel_loop4_inner(data(field11[isub]), data(field21[isub]), r.ldimx, r.ldimy, r.izbeg, stream),
              "kernel_loop4 (Born fields)");
}

// ---- kernel_type2 ----------------------------------------------------------------------------------------------------

void
kernel_type2_base::fwd_main_loop(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    ensure_volumes();
    fwd_main_loop_impl(isub, pml_x, pml_y, pml_z, pml_any);
}
void
kernel_type2_base::fwd_main_loop_2(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    ensure_volumes();
    fwd_main_loop_2_impl(isub, pml_x, pml_y, pml_z, pml_any);
}
void
kernel_type2_base::adj_main_loop(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    ensure_volumes();
    adj_main_loop_impl(isub, pml_x, pml_y, pml_z, pml_any);
}
void
kernel_type2_base::adj_main_loop_2(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    ensure_volumes();
    adj_main_loop_2_impl(isub, pml_x, pml_y, pml_z, pml_any);
}
void
kernel_type2_base::kernel_loop4_derivative(int isub)
{
    ensure_volumes();
    kernel_loop4_derivative_impl(isub);
}
void
kernel_type2_base::snaps_loop(int isub)
{
    ensure_volumes();
    snaps_loop_impl(isub);
}
void
kernel_type2_base::update_grad_loop_1(const int isub, cart_volume<float>* s3, cart_volume<float>* s4,
                                      cart_volume<float>* g, cart_volume<float>* a3, cart_volume<float>* a4)
{
    ensure_volumes();
    update_grad_loop_1_impl(isub, s3, s4, g, a3, a4);
}
void
kernel_type2_base::update_grad_loop_2(const int isub, cart_volume<float>* d1x, cart_volume<float>* d1y,
                                      cart_volume<float>* d1z, cart_volume<float>* d2x, cart_volume<float>* d2y,
                                      cart_volume<float>* d2z, cart_volume<float>* g)
{
    ensure_volumes();
    update_grad_loop_2_impl(isub, d1x, d1y, d1z, d2x, d2y, d2z, g);
}

kernel_type2::kernel_type2(int bc_opt_, int order_, int nsub_, double d1_, double d2_, double d3_, float dt_,
                           cart_volume<float>**& f1, cart_volume<float>**& f2, cart_volume<float>**& f3,
                           cart_volume<float>**& f4, cart_volume<float>**& f3r, cart_volume<float>**& f4r,
                           cart_volume<float>**& m1, cart_volume<float>**& m2, cart_volume<float>**& m3,
                           cart_volume<float>**& m4, cart_volume<float>**& m11, cart_volume<float>**& m5,
                           cart_volume<float>**& m6, cart_volume<float>**& m7, cart_volume<float>**& dx1,
                           cart_volume<float>**& dy1, cart_volume<float>**& dz1, cart_volume<float>**& dx2,
                           cart_volume<float>**& dy2, cart_volume<float>**& dz2,
                           std::vector<cart_volume<float>*>& snap1, std::vector<cart_volume<float>*>& snap2,
                           pml_volumes_cg3d&, const pml_coef3d_gpu&, const sponge_coef3d_gpu& sponge,
                           const bool simple_gradient_)
    : bc_opt(bc_opt_), order(order_), nsub(nsub_), d1(d1_), d2(d2_), d3(d3_), dt(dt_), simple_gradient(simple_gradient_),
      field1_orig(f1), field2_orig(f2), field3_orig(f3), field4_orig(f4), field3_rhs_orig(f3r), field4_rhs_orig(f4r),
      model1_orig(m1), model2_orig(m2), model3_orig(m3), model4_orig(m4), Model11_orig(m11), model5_orig(m5),
      model6_orig(m6), model7_orig(m7), d1x_orig(dx1), d1y_orig(dy1), d1z_orig(dz1), d2x_orig(dx2), d2y_orig(dy2),
      d2z_orig(dz2), snap_field1_orig(snap1), snap_field2_orig(snap2), sponge_coef(sponge)
{
    if (bc_opt != _BDY_SPONGE)
        throw std::runtime_error("kernel_type2: only the sponge boundary (bc_opt = _BDY_SPONGE) is supported");
    if (order != 8)
        throw std::runtime_error("kernel_type2: only FD order 8 is supported");
}

void
kernel_type2::ensure_volumes()
{
    if (cast_done)
        return;
    cast_vols(field1_orig, nsub, field1);
    cast_vols(field2_orig, nsub, field2);
    cast_vols(field3_orig, nsub, field3);
    cast_vols(field4_orig, nsub, field4);
    cast_vols(field3_rhs_orig, nsub, field3_rhs);
    cast_vols(field4_rhs_orig, nsub, field4_rhs);
    cast_vols(model1_orig, nsub, model1);
    cast_vols(model2_orig, nsub, model2);
    cast_vols(model3_orig, nsub, model3);
    cast_vols(model4_orig, nsub, model4);
    cast_vols(Model11_orig, nsub, Model11);
    cast_vols(model5_orig, nsub, model5);
    cast_vols(model6_orig, nsub, model6);
    cast_vols(model7_orig, nsub, model7);
    cast_vols(d1x_orig, nsub, d1x);
    cast_vols(d1y_orig, nsub, d1y);
    cast_vols(d1z_orig, nsub, d1z);
    cast_vols(d2x_orig, nsub, d2x);
    cast_vols(d2y_orig, nsub, d2y);
    cast_vols(d2z_orig, nsub, d2z);
    cast_done = true;
}

void
kernel_type2::derivatives(int isub, cart_volume_regular_gpu* f1, cart_volume_regular_gpu* f2)
{
    // kernels/kernel_type2.cu:146-168 (1/d rounded to float as compute_fd_const does)
    const range r = range_of(f1);
    check(launch_fwd_kernel2_drv(data(f1), data(f2), data(Model11[isub]), data(d1x[isub]), data(d1y[isub]),
                                 data(d1z[isub]), data(d2x[isub]), data(d2y[isub]), data(d2z[isub]), r.ixbeg, r.ixend,
                                 r.iybeg, r.iyend, r.izbeg, r.izend, (float)(1.0 / d1), (float)(1.0 / d2),
                                 (float)(1.0 / d3), r.ldimx, r.ldimy, order, nullptr),
          "launch_fwd_kernel2_drv");
}

void
kernel_type2::fwd_main_loop_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    no_pml(pml_x, pml_y, pml_z, pml_any);
    derivatives(isub, field1[isub], field2[isub]);
}

void
kernel_type2::fwd_main_loop_2_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    // kernels/kernel_type2.cu:180-218. The update is the interior form; a sponge that is not all ones is applied by
    // kernel_loop2 as in the reference.
    no_pml(pml_x, pml_y, pml_z, pml_any);
    const range r = range_of(field3[isub]);
    const auto ab = sponge_coef.getSpongeCoef(isub);
    check(launch_update_rho_fwd_main_loop_2_1(
              data(field3[isub]), data(field4[isub]), data(d1x[isub]), data(d2x[isub]), data(d1y[isub]),
              data(d2y[isub]), data(d1z[isub]), data(d2z[isub]), data(field1[isub]), data(field2[isub]),
              data(Model11[isub]), data(model1[isub]), data(model2[isub]), data(model3[isub]), data(model4[isub]),
              data(model5[isub]), data(model6[isub]), data(model7[isub]), nullptr, nullptr, nullptr, nullptr, nullptr,
              nullptr, nullptr, nullptr, nullptr, ab[0], ab[1], ab[2], r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg,
              r.izend, r.ldimx, r.ldimy, dt, d1, d2, d3, 0, order, 0, 0, 0, nullptr),
          "launch_update_rho_fwd_main_loop_2_1");
}

void
kernel_type2::adj_main_loop_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    // kernels/kernel_type2.cu:221-249
    no_pml(pml_x, pml_y, pml_z, pml_any);
    const range r = range_of(field1[isub]);
    const float invdx = (float)(1.0 / d1), invdy = (float)(1.0 / d2), invdz = (float)(1.0 / d3);
    const float invdxx = (float)(1.0 / (d1 * d1)), invdyy = (float)(1.0 / (d2 * d2)), invdzz = (float)(1.0 / (d3 * d3));
    const float invdxy = (float)(1.0 / (4 * d1 * d2)), invdyz = (float)(1.0 / (4 * d2 * d3)),
                invdzx = (float)(1.0 / (4 * d1 * d3));
    check(launch_adj_kernel2_drv(data(d1x[isub]), data(d2x[isub]), data(d1y[isub]), data(d2y[isub]), data(d1z[isub]),
                                 data(d2z[isub]), data(field3_rhs[isub]), data(field4_rhs[isub]), data(field1[isub]),
                                 data(field2[isub]), data(Model11[isub]), data(model1[isub]), data(model2[isub]),
                                 data(model3[isub]), data(model4[isub]), data(model5[isub]), data(model6[isub]),
                                 data(model7[isub]), r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, invdxx,
                                 invdyy, invdzz, invdxy, invdyz, invdzx, r.ldimx, r.ldimy, order, dt, invdx, invdy,
                                 invdz, nullptr),
          "launch_adj_kernel2_drv");
}

void
kernel_type2::adj_main_loop_2_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    // kernels/kernel_type2.cu:252-276
    const range r = range_of(field1[isub]);
    const auto ab = sponge_coef.getSpongeCoef(isub);
    check(launch_update_adj_kernel_main_3(
              data(field3[isub]), data(field4[isub]), data(d1x[isub]), data(d2x[isub]), data(d1y[isub]),
              data(d2y[isub]), data(d1z[isub]), data(d2z[isub]), data(field3_rhs[isub]), data(field4_rhs[isub]), ab[0],
              ab[1], ab[2], r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, dt, d1, d2, d3,
              pml_x, pml_y, pml_z, order, sponge_coef.use_sponge() ? 1 : 0, nullptr),
          "launch_update_adj_kernel_main_3");
}

void
kernel_type2::kernel_loop4_derivative_impl(int isub)
{
    // kernels/kernel_type2.cu:279-293, restricted like kernel_base::kernel_loop4
    if (!free_surface_.empty() && !free_surface_.at(isub))
        return;
    const range r = range_of(field1[isub]);
    check(kernel_loop4_derivatives(data(d1x[isub]), data(d1y[isub]), data(d1z[isub]), data(d2x[isub]),
                                   data(d2y[isub]), data(d2z[isub]), r.ldimx, r.ldimy, r.izbeg, nullptr),
          "kernel_loop4_derivatives");
}

void
kernel_type2::snaps_loop_impl(int isub)
{
    // kernels/kernel_type2.cu:296-303: forward derivatives of the stored sigma
    if ((size_t)isub >= snap_field1_orig.size() || (size_t)isub >= snap_field2_orig.size())
        throw std::runtime_error("snaps_loop: snapshot volumes are not set");
    derivatives(isub, snap_field1_orig[isub]->as<cart_volume_regular_gpu>(),
                snap_field2_orig[isub]->as<cart_volume_regular_gpu>());
}

void
kernel_type2::update_grad_loop_1_impl(int isub, cart_volume<float>* s3, cart_volume<float>* s4, cart_volume<float>* g,
                                      cart_volume<float>* a3, cart_volume<float>* a4)
{
    // kernels/kernel_type2.cu:306-330: the ranges are field1's, snapshot and gradient use their own layout
    const range r = range_of(field1[isub]);
    auto* sv = s3->as<cart_volume_regular_gpu>();
    auto* av = a3->as<cart_volume_regular_gpu>();
    check(update_Vp_gradient(gdata(s3), gdata(s4), gdata(g), gdata(a3), gdata(a4), r.ixbeg, r.ixend, r.iybeg, r.iyend,
                             r.izbeg, r.izend, sv->ax1()->ntot, sv->ax2()->ntot, av->ax1()->ntot, av->ax2()->ntot,
                             nullptr),
          "update_Vp_gradient");
}

void
kernel_type2::update_grad_loop_2_impl(int isub, cart_volume<float>* fx1, cart_volume<float>* fy1,
                                      cart_volume<float>* fz1, cart_volume<float>* fx2, cart_volume<float>* fy2,
                                      cart_volume<float>* fz2, cart_volume<float>* g)
{
    // kernels/kernel_type2.cu:333-365
    const range r = range_of(field1[isub]);
    auto* gv = g->as<cart_volume_regular_gpu>();
    check(launch_update_rho_gradient_rho(gdata(g), gdata(fx1), gdata(fx2), gdata(fy1), gdata(fy2), gdata(fz1),
                                         gdata(fz2), data(field1[isub]), data(field2[isub]), data(Model11[isub]),
                                         data(model1[isub]), data(model2[isub]), data(model3[isub]),
                                         data(model4[isub]), data(model5[isub]), data(model6[isub]),
                                         data(model7[isub]), r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend,
                                         r.ldimx, r.ldimy, gv->ax1()->ntot, gv->ax2()->ntot, d1, d2, d3, order,
                                         simple_gradient ? 1 : 0, nullptr),
          "launch_update_rho_gradient_rho");
}

// ---- kernel_type1_gpu ---------------------------------------------------------------------------------------------

void
kernel_type1_base::fwd_main_loop_2(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    ensure_volumes();
    fwd_main_loop_2_inner(isub, pml_x, pml_y, pml_z, pml_any);
}
void
kernel_type1_base::adj_main_loop_2(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    ensure_volumes();
    adj_main_loop_2_inner(isub, pml_x, pml_y, pml_z, pml_any);
}
void
kernel_type1_base::update_grad_loop(const int isub, cart_volume<float>* s3, cart_volume<float>* s4,
                                    cart_volume<float>* g, cart_volume<float>* a3, cart_volume<float>* a4)
{
    ensure_volumes();
    update_grad_loop_inner(isub, s3, s4, g, a3, a4);
}

kernel_type1_gpu::kernel_type1_gpu(int bc_opt_, int order_, int nsub_, double d1_, double d2_, double d3_, float dt_,
                                   cart_volume<float>**& f1, cart_volume<float>**& f2, cart_volume<float>**& f3,
                                   cart_volume<float>**& f4, cart_volume<float>**&, cart_volume<float>**&,
                                   cart_volume<float>**& m1, cart_volume<float>**& m2, cart_volume<float>**& m3,
                                   cart_volume<float>**& m4, cart_volume<float>**& m5, cart_volume<float>**& m6,
                                   cart_volume<float>**& m7, pml_volumes_cg3d&, const pml_coef3d_gpu&,
                                   const sponge_coef3d_gpu& sponge)
    : nsub(nsub_), bc_opt(bc_opt_), order(order_), d1(d1_), d2(d2_), d3(d3_), dt(dt_), field1_orig(f1), field2_orig(f2),
      field3_orig(f3), field4_orig(f4), model1_orig(m1), model2_orig(m2), model3_orig(m3), model4_orig(m4),
      model5_orig(m5), model6_orig(m6), model7_orig(m7), sponge_coef(sponge)
{
    if (bc_opt != _BDY_SPONGE)
        throw std::runtime_error("kernel_type1_gpu: only the sponge boundary (bc_opt = _BDY_SPONGE) is supported");
    if (order != 8)
        throw std::runtime_error("kernel_type1_gpu: only FD order 8 is supported");
}

void
kernel_type1_gpu::ensure_volumes()
{
    if (cast_done)
        return;
    cast_vols(field1_orig, nsub, field1);
    cast_vols(field2_orig, nsub, field2);
    cast_vols(field3_orig, nsub, field3);
    cast_vols(field4_orig, nsub, field4);
    cast_vols(model1_orig, nsub, model1);
    cast_vols(model2_orig, nsub, model2);
    cast_vols(model3_orig, nsub, model3);
    cast_vols(model4_orig, nsub, model4);
    cast_vols(model5_orig, nsub, model5);
    cast_vols(model6_orig, nsub, model6);
    cast_vols(model7_orig, nsub, model7);
    cast_done = true;
}

void
kernel_type1_gpu::fwd_main_loop_2_inner(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    // kernels/kernel_type1_gpu.cpp:94-119
    no_pml(pml_x, pml_y, pml_z, pml_any);
    const range r = range_of(field1[isub]);
    const int ldimz = field1[isub]->ax3()->ntot;
    const auto ab = sponge_coef.getSpongeCoef(isub);
    check(fwd_main_loop_2_1_inner_gpu_launch(
              data(field3[isub]), data(field4[isub]), data(field1[isub]), data(field2[isub]), data(model1[isub]),
              data(model2[isub]), data(model3[isub]), data(model4[isub]), data(model5[isub]), data(model6[isub]),
              data(model7[isub]), ab[0], ab[1], ab[2], r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx,
              r.ldimy, ldimz, r.ldimx, r.ldimy, ldimz, dt, d1, d2, d3, order, nullptr, 0,
              sponge_coef.use_sponge() ? 1 : 0, 0),
          "fwd_main_loop_2_1_inner_gpu::launch");
}

void
kernel_type1_gpu::adj_main_loop_2_inner(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any)
{
    // kernels/kernel_type1_gpu.cpp:121-146
    no_pml(pml_x, pml_y, pml_z, pml_any);
    const range r = range_of(field1[isub]);
    const int ldimz = field1[isub]->ax3()->ntot;
    const auto ab = sponge_coef.getSpongeCoef(isub);
    check(adj_main_loop_2_inner_gpu_launch(data(field3[isub]), data(field4[isub]), data(field1[isub]),
                                           data(field2[isub]), data(model1[isub]), data(model2[isub]),
                                           data(model3[isub]), data(model4[isub]), data(model5[isub]),
                                           data(model6[isub]), data(model7[isub]), ab[0], ab[1], ab[2], r.ixbeg,
                                           r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, ldimz, dt,
                                           d1, d2, d3, order, nullptr, sponge_coef.use_sponge() ? 1 : 0, 0),
          "adj_main_loop_2_inner_gpu::launch");
}

void
kernel_type1_gpu::update_grad_loop_inner(int, cart_volume<float>* s3, cart_volume<float>* s4, cart_volume<float>* g,
                                         cart_volume<float>* a3, cart_volume<float>* a4)
{
    // kernels/kernel_type1_gpu.cpp:148-166: the ranges come from the adjoint volume
    auto* av = a3->as<cart_volume_regular_gpu>();
    auto* sv = s3->as<cart_volume_regular_gpu>();
    const range r = range_of(av);
    check(launch_update_grad_loop_kernel(gdata(s3), gdata(s4), gdata(g), gdata(a3), gdata(a4), r.ixbeg, r.ixend,
                                         r.iybeg, r.iyend, r.izbeg, r.izend, sv->ax1()->ntot, sv->ax2()->ntot, r.ldimx,
                                         r.ldimy, nullptr),
          "launch_update_grad_loop_kernel");
}
