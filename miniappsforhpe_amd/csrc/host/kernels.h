// Operator classes of the host side, with the reference's constructors and NVI method names, running the gfx950
// kernels of libtti_hip.so.
//
// Reference: kernels/kernel_abstract_base.h:14-88, kernels/kernel_base.h:13-101 (time update / free surface),
// kernels/kernel_type2_base.h:11-117 + kernels/kernel_type2.h:18-115 (variable-density TTI),
// kernels/kernel_type1_base.h:11-55 + kernels/kernel_type1_gpu.h:9-69 (constant-density TTI),
// utils/sponge_coef3d.h, utils/sponge_coef3d_gpu.h:10-93 (sponge holders).
// PML is outside the path: pml_volumes_cg3d / pml_coef3d_gpu are accepted as empty holders and every pml flag must be
// false. With update_born the time-update loops also run on the Born fields (field11..field41, Model81/Model91), as
// the reference's kernel_base does (kernels/kernel_base.cu:93-364); the adjoint loops have no Born branch there.
#pragma once

#include <array>
#include <vector>

#include "grid.h"

class sponge_coef3d
{
public:
    sponge_coef3d() = default;
    sponge_coef3d(const sponge_coef3d&) = delete;
    sponge_coef3d& operator=(const sponge_coef3d&) = delete;
    // per sub-volume, indexed by array index of the axis (size n + nghost as in the reference)
    std::vector<std::vector<float>> x_A_, x_B_, y_A_, y_B_, z_A_, z_B_;
};

class sponge_coef3d_gpu
{
public:
    sponge_coef3d_gpu() = default;
    explicit sponge_coef3d_gpu(const sponge_coef3d& cpu) { setCoeffs(cpu); }
    ~sponge_coef3d_gpu();
    sponge_coef3d_gpu(const sponge_coef3d_gpu&) = delete;
    sponge_coef3d_gpu& operator=(const sponge_coef3d_gpu&) = delete;

    void setCoeffs(const sponge_coef3d& cpu);
    bool use_sponge() const { return active_coefs; }
    std::array<float2 const*, 3> getSpongeCoef(int isub) const;

    int nsubs = 0;
    std::vector<float2*> x_AB_, y_AB_, z_AB_;

private:
    bool active_coefs = false;
};

struct pml_volumes_cg3d
{}; // PML volumes: not part of the path
struct pml_coef3d_gpu
{};

enum
{
    _BDY_SPONGE = 0
};

class kernel_abstract_base
{
public:
    kernel_abstract_base(int sp_order, int bc_opt) : sp_order_(sp_order), bc_opt_(bc_opt) {}
    virtual ~kernel_abstract_base() = default;

    void kernel_loop1(int isub)
    {
        ensure_volumes();
        kernel_loop1_inner(isub, nullptr);
    }
    void kernel_loop2(int isub)
    {
        ensure_volumes();
        kernel_loop2_inner(isub, nullptr);
    }
    void kernel_loop3(int isub)
    {
        ensure_volumes();
        kernel_loop3_inner(isub, nullptr);
    }
    void kernel_loop4(int isub)
    {
        ensure_volumes();
        kernel_loop4_inner(isub, nullptr);
    }
    void adj_kernel_loop1(int isub)
    {
        ensure_volumes();
        adj_kernel_loop1_inner(isub, nullptr);
    }
    void adj_kernel_loop3(int isub)
    {
        ensure_volumes();
        adj_kernel_loop3_inner(isub, nullptr);
    }

protected:
    virtual void ensure_volumes() = 0;
    virtual void kernel_loop1_inner(int isub, hipStream_t stream);
    virtual void kernel_loop2_inner(int isub, hipStream_t stream) = 0;
    virtual void kernel_loop3_inner(int isub, hipStream_t stream);
    virtual void kernel_loop4_inner(int isub, hipStream_t stream) = 0;
    virtual void adj_kernel_loop1_inner(int isub, hipStream_t stream);
    virtual void adj_kernel_loop3_inner(int isub, hipStream_t stream);
    int sp_order_, bc_opt_;
};

class kernel_base : public kernel_abstract_base
{
public:
    kernel_base(int sp_order, int bc_opt, int nsub, float dt, double d1, double d2, double d3, bool update_born,
                cart_volume<float>**& field1, cart_volume<float>**& field2, cart_volume<float>**& field3,
                cart_volume<float>**& field4, cart_volume<float>**& field11, cart_volume<float>**& field21,
                cart_volume<float>**& field31, cart_volume<float>**& field41, const sponge_coef3d_gpu& sponge_coef,
                std::vector<cart_volume<float>**>& Model8, std::vector<cart_volume<float>**>& Model9,
                std::vector<cart_volume<float>**>& Model81, std::vector<cart_volume<float>**>& Model91,
                std::vector<float>& wsi, std::vector<cart_volume<float>**>& Model10);

    // the free-surface mirror acts on the sub-volumes whose top face is the global top only (the reference's
    // kernel_loop4_inner mirrors every sub-volume's own top face, kernels/kernel_base.cu:164-187)
    void set_free_surface(std::vector<bool> top_face) { free_surface_ = std::move(top_face); }

    static constexpr int nmech = 3;

protected:
    void ensure_volumes() override;
    void kernel_loop1_inner(int isub, hipStream_t stream) override;
    void kernel_loop2_inner(int isub, hipStream_t stream) override;
    void kernel_loop3_inner(int isub, hipStream_t stream) override;
    void kernel_loop4_inner(int isub, hipStream_t stream) override;
    void adj_kernel_loop1_inner(int isub, hipStream_t stream) override;
    void adj_kernel_loop3_inner(int isub, hipStream_t stream) override;

private:
    void ensure_memory_variables();
    struct mech_ptrs
    {
        float* m8[nmech];
        float* m9[nmech];
        const float* m10[nmech];
        float w[nmech];
    };
    mech_ptrs mech(int isub) const;
    mech_ptrs born_mech(int isub) const; // Model81 / Model91 in place of Model8 / Model9

    int nsub;
    float dt;
    bool update_born;
    cart_volume<float>**& field1_orig;
    cart_volume<float>**& field2_orig;
    cart_volume<float>**& field3_orig;
    cart_volume<float>**& field4_orig;
    cart_volume<float>**& field11_orig;
    cart_volume<float>**& field21_orig;
    cart_volume<float>**& field31_orig;
    cart_volume<float>**& field41_orig;
    const sponge_coef3d_gpu& sponge_coef;
    std::vector<cart_volume<float>**>& Model8_orig;
    std::vector<cart_volume<float>**>& Model9_orig;
    std::vector<cart_volume<float>**>& Model81_orig;
    std::vector<cart_volume<float>**>& Model91_orig;
    std::vector<cart_volume<float>**>& Model10_orig;
    std::vector<float>& wsi;
    std::vector<cart_volume_regular_gpu*> field1, field2, field3, field4, field11, field21, field31, field41;
    std::array<std::vector<cart_volume_regular_gpu*>, nmech> Model8, Model9, Model10, Model81, Model91;
    bool mech_cast_done = false;
    std::vector<bool> free_surface_;
    bool cast_done = false;
};

class kernel_type2_base
{
public:
    virtual ~kernel_type2_base() = default;
    void fwd_main_loop(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any);
    void fwd_main_loop_2(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any);
    void adj_main_loop(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any);
    void adj_main_loop_2(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any);
    void kernel_loop4_derivative(int isub);
    void snaps_loop(int isub);
    void update_grad_loop_1(const int isub, cart_volume<float>* snap_x_vol, cart_volume<float>* snap_z_vol,
                            cart_volume<float>* grad_Vp_vol, cart_volume<float>* adj_x_vol,
                            cart_volume<float>* adj_z_vol);
    void update_grad_loop_2(const int isub, cart_volume<float>* fwd_field1_xin_vol,
                            cart_volume<float>* fwd_field1_yin_vol, cart_volume<float>* fwd_field1_zin_vol,
                            cart_volume<float>* fwd_field2_xin_vol, cart_volume<float>* fwd_field2_yin_vol,
                            cart_volume<float>* fwd_field2_zin_vol, cart_volume<float>* grad_Rho_vol);

private:
    virtual void ensure_volumes() = 0;
    virtual void fwd_main_loop_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) = 0;
    virtual void fwd_main_loop_2_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) = 0;
    virtual void adj_main_loop_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) = 0;
    virtual void adj_main_loop_2_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) = 0;
    virtual void kernel_loop4_derivative_impl(int isub) = 0;
    virtual void snaps_loop_impl(int isub) = 0;
    virtual void update_grad_loop_1_impl(int isub, cart_volume<float>*, cart_volume<float>*, cart_volume<float>*,
                                         cart_volume<float>*, cart_volume<float>*) = 0;
    virtual void update_grad_loop_2_impl(int isub, cart_volume<float>*, cart_volume<float>*, cart_volume<float>*,
                                         cart_volume<float>*, cart_volume<float>*, cart_volume<float>*,
                                         cart_volume<float>*) = 0;
};

class kernel_type2 : public kernel_type2_base
{
public:
    kernel_type2(int bc_opt, int order, int nsub, double d1, double d2, double d3, float dt,
                 cart_volume<float>**& field1, cart_volume<float>**& field2, cart_volume<float>**& field3,
                 cart_volume<float>**& field4, cart_volume<float>**& field3_rhs, cart_volume<float>**& field4_rhs,
                 cart_volume<float>**& model1_, cart_volume<float>**& model2_, cart_volume<float>**& model3_,
                 cart_volume<float>**& model4_, cart_volume<float>**& Model11_, cart_volume<float>**& model5_,
                 cart_volume<float>**& model6_, cart_volume<float>**& model7_, cart_volume<float>**& dfield1dx,
                 cart_volume<float>**& dfield1dy, cart_volume<float>**& dfield1dz, cart_volume<float>**& dfield2dx,
                 cart_volume<float>**& dfield2dy, cart_volume<float>**& dfield2dz,
                 std::vector<cart_volume<float>*>& snap_field1, std::vector<cart_volume<float>*>& snap_field2,
                 pml_volumes_cg3d& pml, const pml_coef3d_gpu& pml_coef, const sponge_coef3d_gpu& sponge_coef,
                 const bool simple_gradient = false);

    void set_free_surface(std::vector<bool> top_face) { free_surface_ = std::move(top_face); }

private:
    void ensure_volumes() override;
    void fwd_main_loop_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) override;
    void fwd_main_loop_2_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) override;
    void adj_main_loop_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) override;
    void adj_main_loop_2_impl(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) override;
    void kernel_loop4_derivative_impl(int isub) override;
    void snaps_loop_impl(int isub) override;
    void update_grad_loop_1_impl(int isub, cart_volume<float>*, cart_volume<float>*, cart_volume<float>*,
                                 cart_volume<float>*, cart_volume<float>*) override;
    void update_grad_loop_2_impl(int isub, cart_volume<float>*, cart_volume<float>*, cart_volume<float>*,
                                 cart_volume<float>*, cart_volume<float>*, cart_volume<float>*,
                                 cart_volume<float>*) override;
    void derivatives(int isub, cart_volume_regular_gpu* f1, cart_volume_regular_gpu* f2);

    bool cast_done = false;
    int bc_opt, order, nsub;
    double d1, d2, d3;
    float dt;
    bool simple_gradient;
    cart_volume<float>**& field1_orig;
    cart_volume<float>**& field2_orig;
    cart_volume<float>**& field3_orig;
    cart_volume<float>**& field4_orig;
    cart_volume<float>**& field3_rhs_orig;
    cart_volume<float>**& field4_rhs_orig;
    cart_volume<float>**& model1_orig;
    cart_volume<float>**& model2_orig;
    cart_volume<float>**& model3_orig;
    cart_volume<float>**& model4_orig;
    cart_volume<float>**& Model11_orig;
    cart_volume<float>**& model5_orig;
    cart_volume<float>**& model6_orig;
    cart_volume<float>**& model7_orig;
    cart_volume<float>**& d1x_orig;
    cart_volume<float>**& d1y_orig;
    cart_volume<float>**& d1z_orig;
    cart_volume<float>**& d2x_orig;
    cart_volume<float>**& d2y_orig;
    cart_volume<float>**& d2z_orig;
    std::vector<cart_volume<float>*>& snap_field1_orig;
    std::vector<cart_volume<float>*>& snap_field2_orig;
    const sponge_coef3d_gpu& sponge_coef;
    std::vector<cart_volume_regular_gpu*> field1, field2, field3, field4, field3_rhs, field4_rhs, model1, model2,
        model3, model4, Model11, model5, model6, model7, d1x, d1y, d1z, d2x, d2y, d2z;
    std::vector<bool> free_surface_;
};

class kernel_type1_base
{
public:
    virtual ~kernel_type1_base() = default;
    void fwd_main_loop_2(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any);
    void adj_main_loop_2(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any);
    void update_grad_loop(const int isub, cart_volume<float>* snap_x_vol, cart_volume<float>* snap_z_vol,
                          cart_volume<float>* grad_Vp_vol, cart_volume<float>* adj_x_vol,
                          cart_volume<float>* adj_z_vol);

private:
    virtual void ensure_volumes() = 0;
    virtual void fwd_main_loop_2_inner(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) = 0;
    virtual void adj_main_loop_2_inner(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) = 0;
    virtual void update_grad_loop_inner(int isub, cart_volume<float>*, cart_volume<float>*, cart_volume<float>*,
                                        cart_volume<float>*, cart_volume<float>*) = 0;
};

class kernel_type1_gpu : public kernel_type1_base
{
public:
    kernel_type1_gpu(int bc_opt, int order, int nsub, double d1, double d2, double d3, float dt,
                     cart_volume<float>**& field1, cart_volume<float>**& field2, cart_volume<float>**& field3,
                     cart_volume<float>**& field4, cart_volume<float>**& field3_rhs, cart_volume<float>**& field4_rhs,
                     cart_volume<float>**& model1_, cart_volume<float>**& model2_, cart_volume<float>**& model3_,
                     cart_volume<float>**& model4_, cart_volume<float>**& model5_, cart_volume<float>**& model6_,
                     cart_volume<float>**& model7_, pml_volumes_cg3d& pml, const pml_coef3d_gpu& pml_coef,
                     const sponge_coef3d_gpu& sponge_coef);

private:
    void ensure_volumes() override;
    void fwd_main_loop_2_inner(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) override;
    void adj_main_loop_2_inner(int isub, bool pml_x, bool pml_y, bool pml_z, bool pml_any) override;
    void update_grad_loop_inner(int isub, cart_volume<float>*, cart_volume<float>*, cart_volume<float>*,
                                cart_volume<float>*, cart_volume<float>*) override;

    bool cast_done = false;
    int nsub, bc_opt, order;
    double d1, d2, d3;
    float dt;
    cart_volume<float>**& field1_orig;
    cart_volume<float>**& field2_orig;
    cart_volume<float>**& field3_orig;
    cart_volume<float>**& field4_orig;
    cart_volume<float>**& model1_orig;
    cart_volume<float>**& model2_orig;
    cart_volume<float>**& model3_orig;
    cart_volume<float>**& model4_orig;
    cart_volume<float>**& model5_orig;
    cart_volume<float>**& model6_orig;
    cart_volume<float>**& model7_orig;
    const sponge_coef3d_gpu& sponge_coef;
    std::vector<cart_volume_regular_gpu*> field1, field2, field3, field4, model1, model2, model3, model4, model5,
        model6, model7;
};
