// Variable-density TTI driver on one rank: zones -> sub-volumes, 46 volumes per sub-volume, the forward loop with
// snapshots every xcorr_step steps and the adjoint loop with the imaging conditions, in the step order of the
// reference driver.
//
// Reference: simulator2/simulator2SetUp.h, simulator2/simulator2SetUp.cu:51-417 (constructor), :433-453 (zones),
// :455-475 (init_cijs), :503-515 (kernels), :517-814 (execute), :816-857 (source, pressure).
//
// Decisions where the reference is inconsistent (each changes a result the reference leaves undefined or wrong):
//  * Model11 = 1/rho is initialised (rho defaults to 1); the reference leaves it zero (rho = 1/0).
//  * The free-surface mirrors act only on sub-volumes whose top face is the global top; the reference applies them to
//    every sub-volume's own top face, overwriting exchanged ghosts.
//  * The derivative exchange list holds the six derivatives; the reference lists dfield2dz twice and omits dfield2dx.
//  * The snapshot store holds ceil(niter / xcorr_step) steps (the reference reserves floor and overflows otherwise),
//    and its size sum starts from zero.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "comm.h"
#include "decomp.h"
#include "grid.h"
#include "phase_timer.h"
#include "halo.h"
#include "kernels.h"
#include "snap.h"

class simulator2SetUp
{
public:
    simulator2SetUp() = delete;
    simulator2SetUp(int npoints, int npml, int niter, int xcorr_step, int nzones[3], int ranks[3], int radius_lo,
                    int radius_hi, float bitPerFloat, int fwd_only, MPI_Comm communicator);
    ~simulator2SetUp();
    simulator2SetUp(const simulator2SetUp&) = delete;
    simulator2SetUp& operator=(const simulator2SetUp&) = delete;

    // forward then (unless fwd_only) adjoint; prints the reference's timing summary when verbose
    void execute(bool has_src_p, bool has_qp_);

    // ---- additions for tests and embedding (no counterpart in the reference) ----
    // constant medium; recomputes model1..7 and Model11 of every sub-volume. dt stays the constructor's (the rule of
    // simulator2/simulator2SetUp.cu:127-129 for vp = 1500 m/s and 1 m cells: 2.70e-4 s), and the 8th-order
    // staggered scheme is stable for it while the fastest qP speed (vp sqrt(1 + 2 eps) for eps >= delta) stays
    // below 1 / (dt sqrt(3) sum |c_n|) = 1664 m/s
    void set_medium(float vp, float eps, float delta, float vs, float theta, float phi, float rho);
    void forward(bool has_src_p);
    void adjoint();
    void set_verbose(bool v) { verbose_ = v; }
    // Rebuild the snapshot store before the forward pass: steps beyond max_mem_steps (< 0: the host-RAM budget only)
    // go to a temporary file in the scratch folder
    void set_snapshot_store(const std::string& scratch, int max_mem_steps);
    void set_free_surface(bool on);
    int num_subs() const { return nsubs; }
    float get_dt() const { return dt; }
    cart_volume_regular_gpu* volume(const std::string& name, int isub);
    const decompositionManager3D* decomposition() const { return decompmgr; }
    struct timings
    {
        double forward_s = 0, adjoint_s = 0, velocity_s = 0, stress_s = 0, halo_s = 0;
    };
    const timings& get_timings() const { return times_; }
    // Fused time steps (on by default): with one zone, uncompressed snapshots and no viscoacoustic memory variables,
    // each forward step is one tti_fused_fwd_step and each adjoint step one tti_fused_adj_step (sigma ping-pongs
    // between field1/field2 and a second pair), with the same free surface, the snapshots held in device memory
    // (FP32, the volumes the reference stores) and the imaging conditions of the reference at the snapshot steps.
    // Other cases, or snapshots that do not fit in device memory, run the reference's three-kernel steps. Set
    // before the forward pass; off gives the reference structure everywhere.
    void set_fused(bool on) { fused_ = on; }
    // whether the forward pass runs (or ran) the fused steps
    bool fused_active() const;
    // The kernel_base update_born switch (the reference's driver always passes false, simulator2SetUp.cu:506): on,
    // the time-update loops also advance the Born fields field11..field41 with their memory variables Model81 /
    // Model91 (allocated here, zero, as the reference allocates them, simulator2SetUp.cu:231-255). Set before the
    // forward pass; the steps then take the three-kernel path, whose loops carry the Born fields.
    void set_update_born(bool on);

private:
    int get_n_for_zone(int izone, int nzones) const;
    void init_cijs(int isub);
    void initKernels();
    void initializeSource(float pressure);
    void computefield5(int isub);
    void halo(haloManager3D_gpu* h);
    void allocate_memory_variables();
    void allocate_born();

    MPI_Comm _comm;
    int _rank = 0;
    double d1 = 1.0, d2 = 1.0, d3 = 1.0;
    float dt = 0.002734f;
    float vp = 1500.0f, eps = 0.0f, delta = 0.0f, vs = 0.0f, theta = 0.0f, phi = 0.0f, rho = 1.0f;
    int bc_opt = _BDY_SPONGE;
    int NMECH = 3;
    const int order = 8;

    std::unique_ptr<axis> axgiX, axgiY, axgiZ, axgX, axgY, axgZ;
    std::vector<std::unique_ptr<axis>> local_axes_[3], local_axes_noHalo_[3];

    // per-sub-volume arrays, as the reference's kernels take them
    cart_volume<float>**field1 = nullptr, **field2 = nullptr, **field3 = nullptr, **field4 = nullptr;
    cart_volume<float>**field11 = nullptr, **field21 = nullptr, **field31 = nullptr, **field41 = nullptr;
    cart_volume<float>**field3_rhs = nullptr, **field4_rhs = nullptr;
    cart_volume<float>**model1 = nullptr, **model2 = nullptr, **model3 = nullptr, **model4 = nullptr;
    cart_volume<float>**Model11 = nullptr, **model5 = nullptr, **model6 = nullptr, **model7 = nullptr;
    cart_volume<float>**dfield1dx = nullptr, **dfield1dy = nullptr, **dfield1dz = nullptr;
    cart_volume<float>**dfield2dx = nullptr, **dfield2dy = nullptr, **dfield2dz = nullptr;
    cart_volume<float>**field5_gpu = nullptr, **grad1 = nullptr, **grad2 = nullptr;
    std::vector<cart_volume<float>*> snap_field1, snap_field2, snap_field3, snap_field4;
    std::vector<cart_volume<float>**> Model8{ nullptr, nullptr, nullptr }, Model9{ nullptr, nullptr, nullptr },
        Model81{ nullptr, nullptr, nullptr }, Model91{ nullptr, nullptr, nullptr }, Model10{ nullptr, nullptr, nullptr };
    std::vector<cart_volume<float>**> all_arrays_;
    pml_volumes_cg3d pml;
    const pml_coef3d_gpu pml_coef{};
    kernel_base* kernel_base_ = nullptr;
    kernel_type2* kernel = nullptr;
    std::vector<bool> top_face_;

    std::vector<cart_volume<float>*> corrBuffList;
    std::vector<int> corrBuff_size;
    float memoryPercent = 0.88f;
    int numSnaps = 0;
    std::unique_ptr<rtm_snap> snapReaderWriter;
    hipEvent_t writeSnapshotsCompleteEvent_ = nullptr, readSnapshotsCompleteEvent_ = nullptr;
    std::unique_ptr<file_snap> file_snap_p;
    std::string scratch_folder{ "test_snapshot_area" };
    int next_snap = 0;
    int _xcorr_step;
    float _bitPerFloat;
    bool useZFP_ = false;    // 0 < bitPerFloat < 32: snapshots coded by the fixed-rate ZFP codec
    snapshot_zfp_state zfp_; // its stream, fields and pinned buffers
    // the snapshot store calls of execute, compressed or not (simulator2/simulator2SetUp.cu:609-659, :740-777)
    void snap_lock_write();
    void snap_start_write();
    void snap_start_read();
    void snap_finish_read();

    sponge_coef3d_gpu sponge_coef_gpu;
    sponge_coef3d sponge_coef_cpu;
    std::vector<float> wsi{ 1.0f, 1.0f, 1.0f };

    std::vector<int> numPts_[3];
    int _npml, _npoints, _niter, _fwd_only;
    int radius, r_lo, r_hi;
    int nsubs = 0;
    int n_total[3] = { 0, 0, 0 };
    std::vector<cart_volume<float>**> s_fields_for_halo, sd_fields_for_halo;
    decompositionManager3D* decompmgr = nullptr;
    haloManager3D_gpu* halomgr_s = nullptr;
    haloManager3D_gpu* halomgr_sd = nullptr;

    // fused path (see set_fused)
    bool fused_ = true;
    bool fused_run_ = false;   // the forward pass ran fused; the adjoint follows it
    bool free_surface_ = true; // set_free_surface's state (the fused steps apply it themselves)
    std::unique_ptr<cart_volume_regular_gpu> alt1_, alt2_; // the second sigma pair of the ping-pong
    struct DevSnap
    {
        std::unique_ptr<cart_volume_regular_gpu> f1, f2, f3, f4; // as snap_field1..4 of one snapshot step
    };
    std::vector<DevSnap> dev_snaps_;
    bool fused_possible() const;
    bool fused_snapshots_fit() const;
    void forward_fused(bool has_src_p);
    void adjoint_fused();
    void fused_step(bool forward, float* const s_in[2], float* const s_out[2]);

    bool verbose_ = false;
    bool forward_done_ = false;
    bool qp_ = false; // viscoacoustic memory variables in use (execute's has_qp_)
    bool born_ = false; // kernel_base update_born (set_update_born)
    timings times_;
    tti::PhaseTimer timer_; // phase device times, resolved in batches (no wait per phase)
};
