// Staggered-grid elastic TTI ("eTTI") driver on one rank: zones -> sub-volumes, the 9 wavefield, 22 model, 6 strain
// and 12 mixed-interpolation volumes per sub-volume, the forward loop (velocity, strain, stress, pressure) with
// 9-field snapshots every xcorr_step steps, and the adjoint loop with the velocity and density imaging conditions,
// in the step order of the reference driver.
//
// Reference: simulator3/simulator3SetUp.h, simulator3/simulator3SetUp.cu:302-708 (constructor), :753-766 (zones),
// :768-818 (init_cijs), :2181-2657 (execute), :2659-2708 (source, pressure).
//
// Decisions where the reference is inconsistent (each changes a result the reference leaves undefined or wrong):
//  * The stencil coefficient tables stay zero unless set_coefficients fills them, as in the reference (its
//    init_constant is never called); the kernels then leave every field unchanged, as the reference's do.
//  * The free-surface mirrors act only on sub-volumes whose top face is the global top; the reference applies them to
//    every sub-volume's own top face, zeroing the first interior plane of inner zones and overwriting exchanged ghosts.
//  * The snapshot store holds ceil(niter / xcorr_step) steps and its size sum starts from zero (the reference
//    reserves floor(niter / xcorr_step) steps and sums into an uninitialised counter).
//  * Sub-volume sizes need not be multiples of the reference's 32 x 16 tiles: every kernel bounds its interior.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "comm.h"
#include "decomp.h"
#include "grid.h"
#include "phase_timer.h"
#include "halo.h"
#include "snap.h"

class simulator3SetUp
{
public:
    simulator3SetUp() = delete;
    simulator3SetUp(int npoints, int npml, int niter, int xcorr_step, int nzones[3], int ranks[3], int radius_lo,
                    int radius_hi, float bitPerFloat, int fwd_only, MPI_Comm communicator);
    ~simulator3SetUp();
    simulator3SetUp(const simulator3SetUp&) = delete;
    simulator3SetUp& operator=(const simulator3SetUp&) = delete;

    // forward then (unless fwd_only) adjoint; has_qp_ is accepted and unused, as in the reference
    void execute(bool has_src_p, bool has_qp_);

    // ---- additions for tests and embedding (no counterpart in the reference) ----
    // stencil tables df (staggered first derivative) and si (diagonal interpolation), 4 entries each
    void set_coefficients(const float* df, const float* si);
    void forward(bool has_src_p);
    void adjoint();
    void set_verbose(bool v) { verbose_ = v; }
    // Rebuild the snapshot store before the forward pass: steps beyond max_mem_steps (< 0: the host-RAM budget only)
    // go to a temporary file in the scratch folder
    void set_snapshot_store(const std::string& scratch, int max_mem_steps);
    void set_free_surface(bool on) { free_surface_ = on; }
    int num_subs() const { return nsubs; }
    cart_volume_regular_gpu* volume(const std::string& name, int isub);
    const float* zprime_data(int isub) const { return zprime[isub]; }
    const decompositionManager3D* decomposition() const { return decompmgr; }
    struct timings
    {
        double forward_s = 0, adjoint_s = 0, velocity_s = 0, strain_s = 0, stress_s = 0, halo_s = 0;
    };
    const timings& get_timings() const { return times_; }

private:
    int get_n_for_zone(int izone, int nzones) const;
    void init_cijs(int isub);
    void initializeSource(float p);
    void computeP(int isub);
    void halo(haloManager3D_gpu* h);
    void velocity_all();
    void strain_all();
    void stress_all();
    void mirrors();
    void update_g_v(int isub);
    void update_g_r(int isub);

    MPI_Comm _comm;
    int _rank = 0;
    double d1 = 1.0, d2 = 1.0, d3 = 1.0;

    std::unique_ptr<axis> axgiX, axgiY, axgiZ, axgX, axgY, axgZ;
    std::vector<std::unique_ptr<axis>> local_axes_[3], local_axes_noHalo_[3];

    // per-sub-volume arrays, as the reference's kernels take them
    cart_volume<float>**p_gpu = nullptr, **g1 = nullptr, **g2 = nullptr;
    cart_volume<float>** f[22] = {};  // f[1]..f[21]: velocities f1..f3, stresses f4..f9, mixes f10..f21
    cart_volume<float>** M[23] = {};  // M[1]..M[22]
    cart_volume<float>** d[10] = {};  // d[4]..d[9]
    std::vector<float*> zprime;
    std::vector<cart_volume<float>*> snap_f[10]; // snap_f[1]..snap_f[9]
    std::vector<cart_volume<float>**> all_arrays_;
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

    std::vector<int> numPts_[3];
    int _npml, _npoints, _niter, _fwd_only;
    int radius, r_lo, r_hi;
    int nsubs = 0;
    int n_total[3] = { 0, 0, 0 };
    decompositionManager3D* decompmgr = nullptr;
    haloManager3D_gpu* halomgr_s = nullptr;
    haloManager3D_gpu* halomgr_v = nullptr;
    haloManager3D_gpu* halomgr_vd = nullptr;

    bool verbose_ = false;
    bool free_surface_ = true;
    bool forward_done_ = false;
    timings times_;
    tti::PhaseTimer timer_; // phase device times, resolved in batches (no wait per phase)
};
