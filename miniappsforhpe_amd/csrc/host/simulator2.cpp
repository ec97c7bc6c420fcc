#include "simulator2.h"

#include <chrono>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>

#include "tti_fused.h"
#include "tti_grid.h"
#include "tti_hip.h"

namespace {

cart_volume<float>**
new_array(int n, std::vector<cart_volume<float>**>& owner)
{
    auto** a = new cart_volume<float>*[n]();
    owner.push_back(a);
    return a;
}

} // namespace

simulator2SetUp::simulator2SetUp(int npoints, int npml, int niter, int xcorr_step, int nzones[3], int ranks[3],
                                 int radius_lo, int radius_hi, float bitPerFloat, int fwd_only, MPI_Comm communicator)
    : _comm(communicator), _xcorr_step(xcorr_step), _bitPerFloat(bitPerFloat), _npml(npml), _npoints(npoints),
      _niter(niter), _fwd_only(fwd_only)
{
    if (nzones[0] < 1 || nzones[1] < 1 || nzones[2] < 1 || ranks[0] < 1 || ranks[1] < 1 || ranks[2] < 1)
        throw std::runtime_error("simulator2SetUp: zones and ranks must be >= 1");
    if (npoints < 1 || npml < 0 || niter < 0 || xcorr_step < 1)
        throw std::runtime_error("simulator2SetUp: invalid sizes");
    MPI_Comm_rank(_comm, &_rank);
    r_lo = radius_lo;
    r_hi = radius_hi;
    radius = std::max(radius_lo, radius_hi);
    if (radius != 4)
        throw std::runtime_error("simulator2SetUp: the order-8 stencils need a halo radius of 4");

    // zones: [npoints] | [npml npoints] | [npml npoints ... npoints npml] (simulator2/simulator2SetUp.cu:419-431)
    int* numPts[3];
    for (int d = 0; d < 3; ++d) {
        for (int z = 0; z < nzones[d]; ++z)
            numPts_[d].push_back(get_n_for_zone(z, nzones[d]));
        numPts[d] = numPts_[d].data();
        n_total[d] = 0;
        for (int n : numPts_[d])
            n_total[d] += n;
    }
    decompmgr = new decompositionManager3D(nzones, numPts, nullptr, nullptr, ranks, communicator, 0);

    const float dx = 1.0f, dy = 1.0f, dz = 1.0f;
    axgiX.reset(new axis(0, dx, _npoints));
    axgiY.reset(new axis(0, dy, _npoints));
    axgiZ.reset(new axis(0, dz, _npoints));
    axgX.reset(new axis(axgiX->o - _npml * axgiX->d, axgiX->e + _npml * axgiX->d, axgiX->d));
    axgY.reset(new axis(axgiY->o - _npml * axgiY->d, axgiY->e + _npml * axgiY->d, axgiY->d));
    axgZ.reset(new axis(axgiZ->o - _npml * axgiZ->d, axgiZ->e + _npml * axgiZ->d, axgiZ->d));

    // dt rule of simulator2/simulator2SetUp.cu:127-129
    float dt_l = 1.0f / (1.0f * vp * std::sqrt(1.0f / (dx * dx) + 1.0f / (dy * dy) + 1.0f / (dz * dz)));
    dt_l *= 0.77f;
    dt = 0.91f * dt_l;

    int nsd[3];
    decompmgr->getSplitNumLocalSubDom(nsd);
    const axis* gax[3] = { axgX.get(), axgY.get(), axgZ.get() };
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < nsd[d]; ++s) {
            const int off = decompmgr->getOffset(s, d);
            const float origin = gax[d]->o + float(off) * gax[d]->d;
            const int n = decompmgr->getNumPtsSplit(s, d);
            // x rows padded to 128 bytes (simulator2/simulator2SetUp.cu:140-149)
            local_axes_[d].emplace_back(d == 0 ? new axis(origin, gax[d]->d, n, radius,
                                                          AlignmentElem(AlignMemBytes::CACHELINE, sizeof(float)))
                                               : new axis(origin, gax[d]->d, n, radius));
            local_axes_noHalo_[d].emplace_back(new axis(origin, gax[d]->d, n));
        }
    nsubs = nsd[0] * nsd[1] * nsd[2];

    auto& A = all_arrays_;
    field1 = new_array(nsubs, A);
    field2 = new_array(nsubs, A);
    field3 = new_array(nsubs, A);
    field4 = new_array(nsubs, A);
    field3_rhs = new_array(nsubs, A);
    field4_rhs = new_array(nsubs, A);
    Model11 = new_array(nsubs, A);
    model1 = new_array(nsubs, A);
    model2 = new_array(nsubs, A);
    model3 = new_array(nsubs, A);
    model4 = new_array(nsubs, A);
    model5 = new_array(nsubs, A);
    model6 = new_array(nsubs, A);
    model7 = new_array(nsubs, A);
    dfield1dx = new_array(nsubs, A);
    dfield1dy = new_array(nsubs, A);
    dfield1dz = new_array(nsubs, A);
    dfield2dx = new_array(nsubs, A);
    dfield2dy = new_array(nsubs, A);
    dfield2dz = new_array(nsubs, A);
    field5_gpu = new_array(nsubs, A);
    grad1 = new_array(nsubs, A);
    grad2 = new_array(nsubs, A);
    snap_field1.assign(nsubs, nullptr);
    snap_field2.assign(nsubs, nullptr);
    snap_field3.assign(nsubs, nullptr);
    snap_field4.assign(nsubs, nullptr);
    top_face_.assign(nsubs, false);

    for (int isub = 0; isub < nsubs; ++isub) {
        int sid[3];
        decompmgr->getSplitLocalSubDomID(isub, sid);
        const axis *a1 = local_axes_[0][sid[0]].get(), *a2 = local_axes_[1][sid[1]].get(),
                   *a3 = local_axes_[2][sid[2]].get();
        for (auto** arr : A)
            if (arr != grad1)
                arr[isub] = new cart_volume_regular_gpu(a1, a2, a3, true);
        const axis *n1 = local_axes_noHalo_[0][sid[0]].get(), *n2 = local_axes_noHalo_[1][sid[1]].get(),
                   *n3 = local_axes_noHalo_[2][sid[2]].get();
        grad1[isub] = new cart_volume_regular_gpu(n1, n2, n3, true);
        snap_field3[isub] = new cart_volume_regular_gpu(n1, n2, n3, true);
        snap_field4[isub] = new cart_volume_regular_gpu(n1, n2, n3, true);
        snap_field1[isub] = new cart_volume_regular_gpu(a1, a2, a3, true);
        snap_field2[isub] = new cart_volume_regular_gpu(a1, a2, a3, true);
        corrBuffList.push_back(snap_field3[isub]);
        corrBuffList.push_back(snap_field4[isub]);
        corrBuffList.push_back(snap_field1[isub]);
        corrBuffList.push_back(snap_field2[isub]);
        top_face_[isub] = sid[2] == 0;
        init_cijs(isub);
    }
    // compressed snapshots when 0 < bitPerFloat < 32 (simulator2/simulator2SetUp.cu:285-323): the store then holds
    // each volume's coded words, corrBuff_size[i] of them
    useZFP_ = _bitPerFloat > 0 && _bitPerFloat < 32;
    if (useZFP_)
        zfp_.init(_bitPerFloat, corrBuffList, corrBuff_size);
    else
        for (auto* v : corrBuffList) {
            auto* r = v->as<cart_volume_regular>();
            corrBuff_size.push_back(r->ax1()->nvalid * r->ax2()->nvalid * r->ax3()->nvalid);
        }

    // all-ones sponge of length n + nghost per axis and sub-volume (simulator2/simulator2SetUp.cu:326-356)
    for (auto* v : { &sponge_coef_cpu.x_A_, &sponge_coef_cpu.x_B_, &sponge_coef_cpu.y_A_, &sponge_coef_cpu.y_B_,
                     &sponge_coef_cpu.z_A_, &sponge_coef_cpu.z_B_ })
        v->resize(nsubs);
    for (int isub = 0; isub < nsubs; ++isub) {
        int sid[3];
        decompmgr->getSplitLocalSubDomID(isub, sid);
        const axis *a1 = local_axes_[0][sid[0]].get(), *a2 = local_axes_[1][sid[1]].get(),
                   *a3 = local_axes_[2][sid[2]].get();
        sponge_coef_cpu.x_A_[isub].assign(a1->n + a1->nghost, 1.0f);
        sponge_coef_cpu.x_B_[isub].assign(a1->n + a1->nghost, 1.0f);
        sponge_coef_cpu.y_A_[isub].assign(a2->n + a2->nghost, 1.0f);
        sponge_coef_cpu.y_B_[isub].assign(a2->n + a2->nghost, 1.0f);
        sponge_coef_cpu.z_A_[isub].assign(a3->n + a3->nghost, 1.0f);
        sponge_coef_cpu.z_B_[isub].assign(a3->n + a3->nghost, 1.0f);
    }
    sponge_coef_gpu.setCoeffs(sponge_coef_cpu);

    s_fields_for_halo = { field1, field2 };
    sd_fields_for_halo = { dfield1dx, dfield1dy, dfield1dz, dfield2dx, dfield2dy, dfield2dz };
    std::vector<std::vector<int>> w_s[3][2], w_sd[3][2];
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s) {
            w_s[a][s].assign(s_fields_for_halo.size(), std::vector<int>(nsubs, radius));
            w_sd[a][s].assign(sd_fields_for_halo.size(), std::vector<int>(nsubs, radius));
        }
    std::vector<std::vector<bool>> e_s(s_fields_for_halo.size(), std::vector<bool>(nsubs, true));
    std::vector<std::vector<bool>> e_sd(sd_fields_for_halo.size(), std::vector<bool>(nsubs, true));
    halomgr_s = new haloManager3D_gpu(decompmgr, s_fields_for_halo, w_s, e_s, e_s);
    halomgr_sd = new haloManager3D_gpu(decompmgr, sd_fields_for_halo, w_sd, e_sd, e_sd);

    initKernels();

    numSnaps = _fwd_only ? 0 : (_niter + _xcorr_step - 1) / _xcorr_step;
    long nfloats = 0;
    for (int n : corrBuff_size)
        nfloats += n;
    file_snap_p.reset(file_snap_factory::create(_comm, const_cast<char*>(scratch_folder.c_str()), (uint64_t)nfloats,
                                                memoryPercent, numSnaps, 0));
    tti::hip_check(hipEventCreateWithFlags(&writeSnapshotsCompleteEvent_, hipEventDisableTiming), "event");
    tti::hip_check(hipEventCreateWithFlags(&readSnapshotsCompleteEvent_, hipEventDisableTiming), "event");
    snapReaderWriter.reset(
        new rtm_snap_gpu(corrBuffList, writeSnapshotsCompleteEvent_, readSnapshotsCompleteEvent_, useZFP_));
    tti::hip_check(hipDeviceSynchronize(), "simulator2SetUp: setup");
}

simulator2SetUp::~simulator2SetUp()
{
    (void)hipDeviceSynchronize();
    snapReaderWriter.reset(); // joins the copy thread before its volumes go away
    file_snap_p.reset();
    delete kernel;
    delete kernel_base_;
    delete halomgr_s;
    delete halomgr_sd;
    for (auto** arr : all_arrays_) {
        for (int i = 0; i < nsubs; ++i)
            delete arr[i];
        delete[] arr;
    }
    for (auto* v : { &snap_field1, &snap_field2, &snap_field3, &snap_field4 })
        for (auto* p : *v)
            delete p;
    for (auto e : { writeSnapshotsCompleteEvent_, readSnapshotsCompleteEvent_ })
        if (e)
            (void)hipEventDestroy(e);
    delete decompmgr;
}

int
simulator2SetUp::get_n_for_zone(int izone, int nzones) const
{
    if (nzones == 1)
        return _npoints;
    if (nzones == 2)
        return izone ? _npoints : _npml;
    return (izone == 0 || izone == nzones - 1) ? _npml : _npoints;
}

void
simulator2SetUp::init_cijs(int isub)
{
    // model1..7 and Model11 from the constant medium over the whole allocation (ghosts included)
    auto* m = model1[isub]->as<cart_volume_regular_gpu>();
    const long n = (long)m->nelems();
    float* tmp = nullptr;
    tti::hip_check(hipMalloc((void**)&tmp, sizeof(float) * 7 * n), "init_cijs: hipMalloc");
    const float vals[7] = { vp, eps, delta, vs, theta, phi, rho };
    int rc = 0;
    for (int k = 0; k < 7 && rc == 0; ++k)
        rc = tti_fill(tmp + k * n, n, vals[k], nullptr);
    auto g = [&](cart_volume<float>** a) { return a[isub]->as<cart_volume_regular_gpu>()->getData(); };
    if (rc == 0)
        rc = ::init_cijs(n, tmp, tmp + n, tmp + 2 * n, tmp + 3 * n, tmp + 4 * n, tmp + 5 * n, tmp + 6 * n, g(model1),
                         g(model2), g(model3), g(model4), g(model5), g(model6), g(model7), g(Model11), nullptr);
    const hipError_t s = hipDeviceSynchronize();
    (void)hipFree(tmp);
    tti::abi_check(rc, "init_cijs");
    tti::hip_check(s, "init_cijs");
}

void
simulator2SetUp::set_medium(float vp_, float eps_, float delta_, float vs_, float theta_, float phi_, float rho_)
{
    if (!(vp_ > 0) || !(rho_ > 0) || vs_ < 0 || vs_ >= vp_)
        throw std::runtime_error("set_medium: need vp > vs >= 0 and rho > 0");
    vp = vp_;
    eps = eps_;
    delta = delta_;
    vs = vs_;
    theta = theta_;
    phi = phi_;
    rho = rho_;
    for (int isub = 0; isub < nsubs; ++isub)
        init_cijs(isub);
}

void
simulator2SetUp::initKernels()
{
    kernel_base_ = new kernel_base(order, bc_opt, nsubs, dt, d1, d2, d3, born_, field1, field2, field3, field4,
                                   field11, field21, field31, field41, sponge_coef_gpu, Model8, Model9, Model81,
                                   Model91, wsi, Model10);
    kernel = new kernel_type2(bc_opt, order, nsubs, d1, d2, d3, dt, field1, field2, field3, field4, field3_rhs,
                              field4_rhs, model1, model2, model3, model4, Model11, model5, model6, model7, dfield1dx,
                              dfield1dy, dfield1dz, dfield2dx, dfield2dy, dfield2dz, snap_field1, snap_field2, pml,
                              pml_coef, sponge_coef_gpu);
    set_free_surface(true);
}

void
simulator2SetUp::set_free_surface(bool on)
{
    free_surface_ = on;
    std::vector<bool> faces = on ? top_face_ : std::vector<bool>(nsubs, false);
    kernel_base_->set_free_surface(faces);
    kernel->set_free_surface(faces);
}

void
simulator2SetUp::initializeSource(float pressure)
{
    // point initial condition at the global centre (simulator2/simulator2SetUp.cu:816-840)
    for (int i = 0; i < nsubs; ++i) {
        int st[3], sid[3];
        decompmgr->getSplitLocalSubDomID(i, sid);
        bool inside = true;
        int loc[3];
        for (int d = 0; d < 3; ++d) {
            st[d] = decompmgr->getOffset(sid[d], d);
            const int c = n_total[d] / 2;
            inside = inside && st[d] <= c && c < st[d] + decompmgr->getNumPtsSplit(sid[d], d);
            loc[d] = c - st[d];
        }
        if (!inside)
            continue;
        for (auto** f : { field1, field2 }) {
            auto* v = f[i]->as<cart_volume_regular_gpu>();
            tti::abi_check(tti_fill_box(v->getData(), v->ax1()->ntot, v->ax2()->ntot, v->ax1()->ibeg + loc[0],
                                        v->ax2()->ibeg + loc[1], v->ax3()->ibeg + loc[2], 1, 1, 1, pressure, nullptr),
                           "initializeSource");
        }
    }
}

void
simulator2SetUp::computefield5(int isub)
{
    auto* p = field5_gpu[isub]->as<cart_volume_regular_gpu>();
    tti::abi_check(compute_pressure((long)p->nelems(), field1[isub]->as<cart_volume_regular_gpu>()->getData(),
                                    field2[isub]->as<cart_volume_regular_gpu>()->getData(), p->getData(), nullptr),
                   "compute_pressure");
}

namespace {

template <class F>
double
wall(F&& f)
{
    tti::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    auto t0 = std::chrono::steady_clock::now();
    f();
    tti::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace

void
simulator2SetUp::halo(haloManager3D_gpu* h)
{
    timer_.time(times_.halo_s, [&] {
        h->start_update();
        h->finish_update();
    });
}

// ---- fused path -------------------------------------------------------------------------------------------------

namespace {

struct Range
{
    int ixbeg, ixend, iybeg, iyend, izbeg, izend, ldimx, ldimy;
};

Range
range_of(cart_volume<float>* v)
{
    const auto* r = v->as<cart_volume_regular>();
    return { r->ax1()->ibeg, r->ax1()->iend, r->ax2()->ibeg, r->ax2()->iend,
             r->ax3()->ibeg, r->ax3()->iend, r->ax1()->ntot, r->ax2()->ntot };
}

float*
ptr(cart_volume<float>* v)
{
    return v->as<cart_volume_regular_gpu>()->getData();
}

} // namespace

// The fused steps take one volume with the layout their tiles need (interior x start and row length multiples of 4,
// 16-byte aligned data: the padded x axis of the sub-volumes gives that) and order 8.
bool
simulator2SetUp::fused_possible() const
{
    if (!fused_ || nsubs != 1 || useZFP_ || qp_ || born_ || order != 8)
        return false;
    const Range r = range_of(field1[0]);
    return r.ixbeg % 4 == 0 && r.ldimx % 4 == 0 &&
           reinterpret_cast<uintptr_t>(ptr(field1[0])) % 16 == 0 && reinterpret_cast<uintptr_t>(ptr(Model11[0])) % 16 == 0;
}

// The forward pass's snapshots (numSnaps steps of the four snapshot volumes) must fit in 80% of the free device
// memory, besides the second sigma pair.
bool
simulator2SetUp::fused_snapshots_fit() const
{
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
        return false;
    size_t need = 2 * field1[0]->as<cart_volume_regular_gpu>()->getSize();
    for (auto* v : { snap_field1[0], snap_field2[0], snap_field3[0], snap_field4[0] })
        need += (size_t)numSnaps * v->as<cart_volume_regular_gpu>()->getSize();
    return need <= free_b / 10 * 8;
}

bool
simulator2SetUp::fused_active() const
{
    return forward_done_ ? fused_run_ : fused_possible();
}

// One fused step from s_in to s_out (sigma pairs); u, w are field3/field4. The sponge of this driver is all ones,
// which the fused steps take as no sponge arrays (A = B = 1, the same arithmetic).
void
simulator2SetUp::fused_step(bool fwd, float* const s_in[2], float* const s_out[2])
{
    const Range r = range_of(field1[0]);
    auto m = [&](cart_volume<float>** a) { return ptr(a[0]); };
    const int fs = free_surface_ && top_face_[0] ? 1 : 0;
    int rc;
    if (fwd)
        rc = tti_fused_fwd_step(s_in[0], s_in[1], s_out[0], s_out[1], m(field3), m(field4), m(model1), m(model2),
                                m(model3), m(model4), m(model5), m(model6), m(model7), m(Model11), nullptr, nullptr,
                                nullptr, r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, dt,
                                d1, d2, d3, -1, 0.0f, nullptr, fs, 0, 0, nullptr);
    else
        rc = tti_fused_adj_step(s_in[0], s_in[1], s_out[0], s_out[1], m(field3), m(field4), m(model1), m(model2),
                                m(model3), m(model4), m(model5), m(model6), m(model7), m(Model11), nullptr, nullptr,
                                nullptr, r.ixbeg, r.ixend, r.iybeg, r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, dt,
                                d1, d2, d3, nullptr, nullptr, fs, 0, 0, nullptr);
    tti::abi_check(rc, fwd ? "tti_fused_fwd_step" : "tti_fused_adj_step");
}

// Forward pass on the fused steps (the sequence of forward() below, one kernel per step): kernel_loop2 (+ mirror) of
// step 0 primes sigma, then step n reads sigma_n and writes sigma_{n+1} = A sigma_n + B dt u_{n+1} into the other
// pair. At the snapshot steps the volumes the reference stores (u, w interiors; sigma_n with its ghosts) are copied
// into device slots. At the end field1/field2 hold sigma of the last step and field5 its pressure, as in the reference
// (which recomputes field5 every step; only the last one is observable, so it is computed once).
void
simulator2SetUp::forward_fused(bool has_src_p)
{
    if (has_src_p)
        initializeSource(500.0f);
    next_snap = 0;
    auto* f1 = field1[0]->as<cart_volume_regular_gpu>();
    const auto* fr = f1->as<cart_volume_regular>();
    if (!alt1_) {
        alt1_.reset(new cart_volume_regular_gpu(fr->ax1(), fr->ax2(), fr->ax3(), true));
        alt2_.reset(new cart_volume_regular_gpu(fr->ax1(), fr->ax2(), fr->ax3(), true));
    }
    if (!_fwd_only && dev_snaps_.empty()) {
        dev_snaps_.resize(numSnaps);
        for (auto& ds : dev_snaps_) {
            auto mk = [](cart_volume<float>* like) {
                const auto* v = like->as<cart_volume_regular>();
                return std::unique_ptr<cart_volume_regular_gpu>(
                    new cart_volume_regular_gpu(v->ax1(), v->ax2(), v->ax3(), false));
            };
            ds.f1 = mk(snap_field1[0]);
            ds.f2 = mk(snap_field2[0]);
            ds.f3 = mk(snap_field3[0]);
            ds.f4 = mk(snap_field4[0]);
        }
    }
    cart_volume<float>* pair[2][2] = { { field1[0], field2[0] }, { alt1_.get(), alt2_.get() } };
    int cur = 0;
    kernel_base_->kernel_loop2(0);
    kernel_base_->kernel_loop4(0);
    for (int iter = 0; iter < _niter; ++iter) {
        float* const in[2] = { ptr(pair[cur][0]), ptr(pair[cur][1]) };
        float* const out[2] = { ptr(pair[cur ^ 1][0]), ptr(pair[cur ^ 1][1]) };
        timer_.time(times_.velocity_s, [&] { fused_step(true, in, out); });
        if (!_fwd_only && iter == next_snap) {
            DevSnap& ds = dev_snaps_.at(iter / _xcorr_step);
            ds.f3->copyData(field3[0], true);
            ds.f4->copyData(field4[0], true);
            ds.f1->copyData(pair[cur][0], false);
            ds.f2->copyData(pair[cur][1], false);
            next_snap += _xcorr_step;
        }
        cur ^= 1;
    }
    // sigma of the last step (the pair the last step read) into field1/field2, and its pressure
    if (_niter > 0 && (cur ^ 1) == 1) {
        f1->copyData(alt1_.get(), false);
        field2[0]->as<cart_volume_regular_gpu>()->copyData(alt2_.get(), false);
    }
    if (_niter > 0)
        computefield5(0);
    tti::hip_check(hipDeviceSynchronize(), "simulator2SetUp::forward (fused)");
    timer_.flush();
}

// Adjoint pass on the fused steps (the sequence of adjoint() below): step n reads sigma and writes
// A sigma + B dt u' into the other pair; at the snapshot steps the reference's imaging conditions run on the device
// slot of that step: grad1 += u' u_snap + w' w_snap (update_grad_loop_1), and the density image from the derivatives
// of the stored sigma (snaps_loop, derivative mirror) and the adjoint sigma after the step (update_grad_loop_2).
void
simulator2SetUp::adjoint_fused()
{
    next_snap -= _xcorr_step;
    cart_volume<float>* pair[2][2] = { { field1[0], field2[0] }, { alt1_.get(), alt2_.get() } };
    int cur = 0;
    const Range r = range_of(field1[0]);
    for (int iter = _niter - 1; iter >= 0; --iter) {
        float* const in[2] = { ptr(pair[cur][0]), ptr(pair[cur][1]) };
        float* const out[2] = { ptr(pair[cur ^ 1][0]), ptr(pair[cur ^ 1][1]) };
        timer_.time(times_.stress_s, [&] { fused_step(false, in, out); });
        cur ^= 1;
        if (iter == next_snap) {
            DevSnap& ds = dev_snaps_.at(iter / _xcorr_step);
            kernel->update_grad_loop_1(0, ds.f3.get(), ds.f4.get(), grad1[0], field3[0], field4[0]);
            {
                // snaps_loop differentiates the driver's snapshot sigma: point it at the slot for this call (restored
                // on every exit, so the driver's own volumes are never left behind a slot it does not own)
                struct Swap
                {
                    std::vector<cart_volume<float>*>&a, &b;
                    cart_volume<float>*ka, *kb;
                    ~Swap()
                    {
                        a[0] = ka;
                        b[0] = kb;
                    }
                } guard{ snap_field1, snap_field2, snap_field1[0], snap_field2[0] };
                snap_field1[0] = ds.f1.get();
                snap_field2[0] = ds.f2.get();
                kernel->snaps_loop(0);
            }
            halo(halomgr_sd);
            kernel->kernel_loop4_derivative(0);
            // update_grad_loop_2 with the adjoint sigma of this step (the pair just written)
            auto* g = grad2[0]->as<cart_volume_regular_gpu>();
            auto m = [&](cart_volume<float>** a) { return ptr(a[0]); };
            tti::abi_check(launch_update_rho_gradient_rho(
                               g->getData(), m(dfield1dx), m(dfield2dx), m(dfield1dy), m(dfield2dy), m(dfield1dz),
                               m(dfield2dz), ptr(pair[cur][0]), ptr(pair[cur][1]), m(Model11), m(model1), m(model2),
                               m(model3), m(model4), m(model5), m(model6), m(model7), r.ixbeg, r.ixend, r.iybeg,
                               r.iyend, r.izbeg, r.izend, r.ldimx, r.ldimy, g->ax1()->ntot, g->ax2()->ntot, d1, d2, d3,
                               order, 0, nullptr),
                           "launch_update_rho_gradient_rho");
            next_snap -= _xcorr_step;
        }
    }
    if (cur == 1) { // the final adjoint sigma into field1/field2
        field1[0]->as<cart_volume_regular_gpu>()->copyData(alt1_.get(), false);
        field2[0]->as<cart_volume_regular_gpu>()->copyData(alt2_.get(), false);
    }
    tti::hip_check(hipDeviceSynchronize(), "simulator2SetUp::adjoint (fused)");
    timer_.flush();
}

void
simulator2SetUp::forward(bool has_src_p)
{
    // simulator2/simulator2SetUp.cu:525-649
    if (forward_done_)
        throw std::runtime_error("simulator2SetUp: the forward pass runs once per object");
    forward_done_ = true;
    fused_run_ = fused_possible() && (_fwd_only || fused_snapshots_fit());
    if (fused_run_) {
        forward_fused(has_src_p);
        return;
    }
    if (has_src_p)
        initializeSource(500.0f);
    next_snap = 0;
    for (int iter = 0; iter < _niter; ++iter) {
        if (qp_) {
            // viscoacoustic: memory-variable sigma update, then the memory variables (simulator2SetUp.cu:530-545)
            for (int isub = 0; isub < nsubs; ++isub)
                kernel_base_->kernel_loop1(isub);
            for (int isub = 0; isub < nsubs; ++isub)
                kernel_base_->kernel_loop3(isub);
        } else {
            for (int isub = 0; isub < nsubs; ++isub)
                kernel_base_->kernel_loop2(isub);
        }
        halo(halomgr_s);
        timer_.time(times_.velocity_s, [&] {
            for (int isub = 0; isub < nsubs; ++isub)
                kernel_base_->kernel_loop4(isub);
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->fwd_main_loop(isub, false, false, false, false);
        });
        halo(halomgr_sd);
        timer_.time(times_.stress_s, [&] {
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->kernel_loop4_derivative(isub);
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->fwd_main_loop_2(isub, false, false, false, false);
        });
        if (_fwd_only)
            continue;
        if (iter == next_snap) {
            snap_lock_write();
            for (int isub = 0; isub < nsubs; ++isub) {
                snap_field3[isub]->as<cart_volume_regular_gpu>()->copyData(field3[isub], true);
                snap_field4[isub]->as<cart_volume_regular_gpu>()->copyData(field4[isub], true);
            }
            halo(halomgr_s);
            for (int isub = 0; isub < nsubs; ++isub) {
                snap_field1[isub]->as<cart_volume_regular_gpu>()->copyData(field1[isub], false);
                snap_field2[isub]->as<cart_volume_regular_gpu>()->copyData(field2[isub], false);
            }
            tti::hip_check(hipEventRecord(writeSnapshotsCompleteEvent_, nullptr), "hipEventRecord");
            snap_start_write();
            next_snap += _xcorr_step;
        }
        for (int isub = 0; isub < nsubs; ++isub)
            computefield5(isub);
    }
}

void
simulator2SetUp::adjoint()
{
    // simulator2/simulator2SetUp.cu:652-779
    if (!forward_done_ || _fwd_only)
        throw std::runtime_error("simulator2SetUp: the adjoint pass needs the forward snapshots");
    if (fused_run_) {
        adjoint_fused();
        return;
    }
    snap_start_read();
    next_snap -= _xcorr_step;
    for (int iter = _niter - 1; iter >= 0; --iter) {
        timer_.time(times_.stress_s, [&] {
            if (qp_)
                for (int isub = 0; isub < nsubs; ++isub)
                    kernel_base_->adj_kernel_loop3(isub); // simulator2SetUp.cu:669-673
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->adj_main_loop(isub, false, false, false, false);
        });
        halo(halomgr_sd);
        timer_.time(times_.velocity_s, [&] {
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->kernel_loop4_derivative(isub);
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->adj_main_loop_2(isub, false, false, false, false);
            for (int isub = 0; isub < nsubs; ++isub) {
                if (qp_)
                    kernel_base_->adj_kernel_loop1(isub); // simulator2SetUp.cu:710-718
                else
                    kernel_base_->kernel_loop2(isub);
            }
        });
        halo(halomgr_s);
        for (int isub = 0; isub < nsubs; ++isub)
            kernel_base_->kernel_loop4(isub);
        if (iter == next_snap) {
            snap_finish_read();
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->update_grad_loop_1(isub, snap_field3[isub], snap_field4[isub], grad1[isub], field3[isub],
                                           field4[isub]);
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->snaps_loop(isub);
            halo(halomgr_sd);
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->kernel_loop4_derivative(isub);
            for (int isub = 0; isub < nsubs; ++isub)
                kernel->update_grad_loop_2(isub, dfield1dx[isub], dfield1dy[isub], dfield1dz[isub], dfield2dx[isub],
                                           dfield2dy[isub], dfield2dz[isub], grad2[isub]);
            tti::hip_check(hipEventRecord(readSnapshotsCompleteEvent_, nullptr), "hipEventRecord");
            next_snap -= _xcorr_step;
            if (next_snap >= 0)
                snap_start_read();
        }
    }
    tti::hip_check(hipDeviceSynchronize(), "simulator2SetUp::adjoint");
}

void
simulator2SetUp::execute(bool has_src_p, bool has_qp_)
{
    if (has_qp_ && !qp_) {
        if (forward_done_)
            throw std::runtime_error("simulator2SetUp: has_qp_ must be set on the first execute");
        allocate_memory_variables();
        qp_ = true;
    }
    times_.forward_s += wall([&] { forward(has_src_p); });
    timer_.flush();
    if (!_fwd_only)
        times_.adjoint_s += wall([&] { adjoint(); });
        timer_.flush();
    if (verbose_ && _rank == 0) {
        std::printf("Total Time: %g\n", times_.forward_s + times_.adjoint_s);
        std::printf("Max time for Forward: %g\n", times_.forward_s);
        if (!_fwd_only)
            std::printf("Max time for Adjoint: %g\n", times_.adjoint_s);
        std::printf("Max time for velocity: %g\n", times_.velocity_s);
        std::printf("Max time for stress: %g\n", times_.stress_s);
        std::printf("Max time for halo-exchange: %g\n", times_.halo_s);
    }
}

void
simulator2SetUp::set_update_born(bool on)
{
    if (forward_done_)
        throw std::runtime_error("simulator2SetUp: update_born is set before the forward pass");
    if (on == born_)
        return;
    if (on && !field11)
        allocate_born();
    born_ = on;
    // the operator takes update_born at construction (kernels/kernel_base.h:16), so it is built again
    delete kernel_base_;
    kernel_base_ = nullptr;
    kernel_base_ = new kernel_base(order, bc_opt, nsubs, dt, d1, d2, d3, born_, field1, field2, field3, field4,
                                   field11, field21, field31, field41, sponge_coef_gpu, Model8, Model9, Model81,
                                   Model91, wsi, Model10);
    set_free_surface(free_surface_);
}

// The Born fields and their memory variables, zero (simulator2/simulator2SetUp.cu:231-234, :254-255)
void
simulator2SetUp::allocate_born()
{
    auto zero_like_field1 = [this](cart_volume<float>** arr) {
        for (int isub = 0; isub < nsubs; ++isub) {
            const auto* f = field1[isub]->as<cart_volume_regular>();
            arr[isub] = new cart_volume_regular_gpu(f->ax1(), f->ax2(), f->ax3(), true);
        }
    };
    for (auto** arr : { &field11, &field21, &field31, &field41 }) {
        *arr = new_array(nsubs, all_arrays_);
        zero_like_field1(*arr);
    }
    for (int n = 0; n < 3; ++n)
        for (auto* arr : { &Model81[n], &Model91[n] }) {
            *arr = new_array(nsubs, all_arrays_);
            zero_like_field1(*arr);
        }
    tti::hip_check(hipDeviceSynchronize(), "simulator2SetUp: Born fields");
}

// Memory variables of the viscoacoustic loops: NMECH = 3 of each of Model8, Model9 (of sigma_H, sigma_V) and Model10
// (alpha), all ones as the reference initialises them (simulator2/simulator2SetUp.cu:252-262), omegas wsi = 1. They
// are allocated on the first has_qp_ execute rather than in the constructor, so drivers that never use them do not
// hold 9 more volumes per sub-volume. The Born copies (Model81, Model91) come with set_update_born.
void
simulator2SetUp::allocate_memory_variables()
{
    for (int n = 0; n < 3; ++n) {
        for (auto* arr : { &Model8[n], &Model9[n], &Model10[n] }) {
            *arr = new_array(nsubs, all_arrays_);
            for (int isub = 0; isub < nsubs; ++isub) {
                const auto* f = field1[isub]->as<cart_volume_regular>();
                auto* v = new cart_volume_regular_gpu(f->ax1(), f->ax2(), f->ax3(), true);
                v->set_constant(1.0f, false, false);
                (*arr)[isub] = v;
            }
        }
    }
    tti::hip_check(hipDeviceSynchronize(), "simulator2SetUp: memory variables");
}

cart_volume_regular_gpu*
simulator2SetUp::volume(const std::string& name, int isub)
{
    if (isub < 0 || isub >= nsubs)
        throw std::runtime_error("volume: sub-volume index out of range");
    const std::map<std::string, cart_volume<float>**> arrays = {
        { "field1", field1 },       { "field2", field2 },       { "field3", field3 },
        { "field4", field4 },       { "field3_rhs", field3_rhs }, { "field4_rhs", field4_rhs },
        { "model1", model1 },       { "model2", model2 },       { "model3", model3 },
        { "model4", model4 },       { "model5", model5 },       { "model6", model6 },
        { "model7", model7 },       { "Model11", Model11 },     { "dfield1dx", dfield1dx },
        { "dfield1dy", dfield1dy }, { "dfield1dz", dfield1dz }, { "dfield2dx", dfield2dx },
        { "dfield2dy", dfield2dy }, { "dfield2dz", dfield2dz }, { "field5", field5_gpu },
        { "grad1", grad1 },         { "grad2", grad2 },
    };
    auto it = arrays.find(name);
    if (it != arrays.end())
        return it->second[isub]->as<cart_volume_regular_gpu>();
    // Born fields "field11" .. "field41" and memory variables "Model81_0" .. "Model91_2" (set_update_born)
    for (auto& [bname, arr] : { std::pair<const char*, cart_volume<float>**>{ "field11", field11 },
                                { "field21", field21 },
                                { "field31", field31 },
                                { "field41", field41 } })
        if (name == bname) {
            if (!arr)
                throw std::runtime_error("volume: " + name + " is allocated by set_update_born");
            return arr[isub]->as<cart_volume_regular_gpu>();
        }
    for (int n = 0; n < 3; ++n)
        for (auto& [prefix, vec] : { std::pair<const char*, std::vector<cart_volume<float>**>*>{ "Model81_", &Model81 },
                                     { "Model91_", &Model91 } })
            if (name == prefix + std::to_string(n)) {
                if (!(*vec)[n])
                    throw std::runtime_error("volume: " + name + " is allocated by set_update_born");
                return (*vec)[n][isub]->as<cart_volume_regular_gpu>();
            }
    // memory variables "Model8_0" .. "Model10_2" (present once a has_qp_ execute allocated them)
    for (int n = 0; n < 3; ++n)
        for (auto& [prefix, vec] : { std::pair<const char*, std::vector<cart_volume<float>**>*>{ "Model8_", &Model8 },
                                     { "Model9_", &Model9 },
                                     { "Model10_", &Model10 } })
            if (name == prefix + std::to_string(n)) {
                if (!(*vec)[n])
                    throw std::runtime_error("volume: " + name + " is allocated by the first has_qp_ execute");
                return (*vec)[n][isub]->as<cart_volume_regular_gpu>();
            }
    const std::map<std::string, std::vector<cart_volume<float>*>*> lists = {
        { "snap_field1", &snap_field1 },
        { "snap_field2", &snap_field2 },
        { "snap_field3", &snap_field3 },
        { "snap_field4", &snap_field4 },
    };
    auto jt = lists.find(name);
    if (jt != lists.end())
        return (*jt->second)[isub]->as<cart_volume_regular_gpu>();
    throw std::runtime_error("volume: unknown name " + name);
}

void
simulator2SetUp::set_snapshot_store(const std::string& scratch, int max_mem_steps)
{
    if (forward_done_)
        throw std::runtime_error("simulator2SetUp: the snapshot store is set before the forward pass");
    long nfloats = 0;
    for (int n : corrBuff_size)
        nfloats += n;
    scratch_folder = scratch.empty() ? std::string(".") : scratch;
    file_snap_p.reset(); // the old store (and its scratch file) goes first
    file_snap_p.reset(file_snap_factory::create(_comm, const_cast<char*>(scratch_folder.c_str()), (uint64_t)nfloats,
                                                memoryPercent, numSnaps, 0, max_mem_steps));
}

void
simulator2SetUp::snap_lock_write()
{
    if (useZFP_)
        snapReaderWriter->lock_for_write_compressed_fwi_snapshot();
    else
        snapReaderWriter->lock_for_write_uncompressed_fwi_snapshot();
}

void
simulator2SetUp::snap_start_write()
{
    if (useZFP_)
        snapReaderWriter->start_write_compressed_fwi_snapshot(corrBuffList, corrBuff_size, file_snap_p.get(),
                                                              zfp_.zipped, zfp_.fields, zfp_.stream);
    else
        snapReaderWriter->start_write_uncompressed_fwi_snapshot(corrBuffList, corrBuff_size, file_snap_p.get());
}

void
simulator2SetUp::snap_start_read()
{
    if (useZFP_)
        snapReaderWriter->start_read_compressed_fwi_snapshot(corrBuffList, corrBuff_size, file_snap_p.get(),
                                                             zfp_.zipped, zfp_.fields, zfp_.stream);
    else
        snapReaderWriter->start_read_uncompressed_fwi_snapshot(corrBuffList, corrBuff_size, file_snap_p.get());
}

void
simulator2SetUp::snap_finish_read()
{
    if (useZFP_)
        snapReaderWriter->finish_read_compressed_fwi_snapshot();
    else
        snapReaderWriter->finish_read_uncompressed_fwi_snapshot();
}
