#include "simulator3.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>

#include "tti_etti.h"
#include "tti_grid.h"

namespace {

cart_volume<float>**
new_array(int n, std::vector<cart_volume<float>**>& owner)
{
    auto** a = new cart_volume<float>*[n]();
    owner.push_back(a);
    return a;
}

float*
data(cart_volume<float>** a, int isub)
{
    return a[isub]->as<cart_volume_regular_gpu>()->getData();
}

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

simulator3SetUp::simulator3SetUp(int npoints, int npml, int niter, int xcorr_step, int nzones[3], int ranks[3],
                                 int radius_lo, int radius_hi, float bitPerFloat, int fwd_only, MPI_Comm communicator)
    : _comm(communicator), _xcorr_step(xcorr_step), _bitPerFloat(bitPerFloat), _npml(npml), _npoints(npoints),
      _niter(niter), _fwd_only(fwd_only)
{
    if (nzones[0] < 1 || nzones[1] < 1 || nzones[2] < 1 || ranks[0] < 1 || ranks[1] < 1 || ranks[2] < 1)
        throw std::runtime_error("simulator3SetUp: zones and ranks must be >= 1");
    if (npoints < 1 || npml < 0 || niter < 0 || xcorr_step < 1)
        throw std::runtime_error("simulator3SetUp: invalid sizes");
    MPI_Comm_rank(_comm, &_rank);
    r_lo = radius_lo;
    r_hi = radius_hi;
    radius = std::max(radius_lo, radius_hi);
    if (radius != 4)
        throw std::runtime_error("simulator3SetUp: the order-8 stencils need a halo radius of 4");

    // zones: [npoints] | [npml npoints] | [npml npoints ... npoints npml] (simulator3/simulator3SetUp.cu:753-766)
    int* numPts[3];
    for (int a = 0; a < 3; ++a) {
        for (int z = 0; z < nzones[a]; ++z)
            numPts_[a].push_back(get_n_for_zone(z, nzones[a]));
        numPts[a] = numPts_[a].data();
        for (int n : numPts_[a])
            n_total[a] += n;
        if (*std::min_element(numPts_[a].begin(), numPts_[a].end()) < 1)
            throw std::runtime_error("simulator3SetUp: every zone needs at least one point (npml >= 1 with zones > 1)");
    }
    decompmgr = new decompositionManager3D(nzones, numPts, nullptr, nullptr, ranks, communicator, 0);

    const float dx = 1.0f, dy = 1.0f, dz = 1.0f;
    axgiX.reset(new axis(0, dx, _npoints));
    axgiY.reset(new axis(0, dy, _npoints));
    axgiZ.reset(new axis(0, dz, _npoints));
    axgX.reset(new axis(axgiX->o - _npml * axgiX->d, axgiX->e + _npml * axgiX->d, axgiX->d));
    axgY.reset(new axis(axgiY->o - _npml * axgiY->d, axgiY->e + _npml * axgiY->d, axgiY->d));
    axgZ.reset(new axis(axgiZ->o - _npml * axgiZ->d, axgiZ->e + _npml * axgiZ->d, axgiZ->d));

    int nsd[3];
    decompmgr->getSplitNumLocalSubDom(nsd);
    const axis* gax[3] = { axgX.get(), axgY.get(), axgZ.get() };
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < nsd[a]; ++s) {
            const int off = decompmgr->getOffset(s, a);
            const float origin = gax[a]->o + float(off) * gax[a]->d;
            const int n = decompmgr->getNumPtsSplit(s, a);
            // x rows padded to 128 bytes (simulator3/simulator3SetUp.cu:386-395)
            local_axes_[a].emplace_back(a == 0 ? new axis(origin, gax[a]->d, n, radius,
                                                          AlignmentElem(AlignMemBytes::CACHELINE, sizeof(float)))
                                               : new axis(origin, gax[a]->d, n, radius));
            local_axes_noHalo_[a].emplace_back(new axis(origin, gax[a]->d, n));
        }
    nsubs = nsd[0] * nsd[1] * nsd[2];

    // volumes of simulator3/simulator3SetUp.cu:419-547
    auto& A = all_arrays_;
    p_gpu = new_array(nsubs, A);
    g2 = new_array(nsubs, A);
    for (int k = 1; k <= 21; ++k)
        f[k] = new_array(nsubs, A);
    for (int k = 1; k <= 22; ++k)
        M[k] = new_array(nsubs, A);
    for (int k = 4; k <= 9; ++k)
        d[k] = new_array(nsubs, A);
    g1 = new_array(nsubs, A);
    zprime.assign(nsubs, nullptr);
    for (int k = 1; k <= 9; ++k)
        snap_f[k].assign(nsubs, nullptr);
    top_face_.assign(nsubs, false);

    for (int isub = 0; isub < nsubs; ++isub) {
        int sid[3];
        decompmgr->getSplitLocalSubDomID(isub, sid);
        const axis *a1 = local_axes_[0][sid[0]].get(), *a2 = local_axes_[1][sid[1]].get(),
                   *a3 = local_axes_[2][sid[2]].get();
        const axis *n1 = local_axes_noHalo_[0][sid[0]].get(), *n2 = local_axes_noHalo_[1][sid[1]].get(),
                   *n3 = local_axes_noHalo_[2][sid[2]].get();
        for (auto** arr : A)
            arr[isub] = arr == g1 ? new cart_volume_regular_gpu(n1, n2, n3, true)
                                  : new cart_volume_regular_gpu(a1, a2, a3, true);
        tti::hip_check(hipMalloc((void**)&zprime[isub], sizeof(float) * a3->ntot), "simulator3SetUp: zprime");
        top_face_[isub] = sid[2] == 0;
        init_cijs(isub);
        // snapshot buffers: velocities without ghosts, stresses with (simulator3/simulator3SetUp.cu:553-576)
        for (int k = 1; k <= 9; ++k) {
            snap_f[k][isub] = k <= 3 ? new cart_volume_regular_gpu(n1, n2, n3, true)
                                     : new cart_volume_regular_gpu(a1, a2, a3, true);
            corrBuffList.push_back(snap_f[k][isub]);
        }
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

    // halo managers: stresses (s), velocities (v), strains (vd) (simulator3/simulator3SetUp.cu:616-680)
    auto make_halo = [&](std::vector<cart_volume<float>**> fields) {
        std::vector<std::vector<int>> w[3][2];
        for (int a = 0; a < 3; ++a)
            for (int s = 0; s < 2; ++s)
                w[a][s].assign(fields.size(), std::vector<int>(nsubs, radius));
        std::vector<std::vector<bool>> e(fields.size(), std::vector<bool>(nsubs, true));
        return new haloManager3D_gpu(decompmgr, fields, w, e, e);
    };
    halomgr_s = make_halo({ f[4], f[5], f[6], f[7], f[8], f[9] });
    halomgr_v = make_halo({ f[1], f[2], f[3] });
    halomgr_vd = make_halo({ d[4], d[5], d[6], d[7], d[8], d[9] });

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
    tti::hip_check(hipDeviceSynchronize(), "simulator3SetUp: setup");
}

simulator3SetUp::~simulator3SetUp()
{
    (void)hipDeviceSynchronize();
    snapReaderWriter.reset(); // joins the copy thread before its volumes go away
    file_snap_p.reset();
    delete halomgr_s;
    delete halomgr_v;
    delete halomgr_vd;
    for (auto** arr : all_arrays_) {
        for (int i = 0; i < nsubs; ++i)
            delete arr[i];
        delete[] arr;
    }
    for (int k = 1; k <= 9; ++k)
        for (auto* p : snap_f[k])
            delete p;
    for (float* z : zprime)
        if (z)
            (void)hipFree(z);
    for (auto e : { writeSnapshotsCompleteEvent_, readSnapshotsCompleteEvent_ })
        if (e)
            (void)hipEventDestroy(e);
    delete decompmgr;
}

int
simulator3SetUp::get_n_for_zone(int izone, int nzones) const
{
    if (nzones == 1)
        return _npoints;
    if (nzones == 2)
        return izone ? _npoints : _npml;
    return (izone == 0 || izone == nzones - 1) ? _npml : _npoints;
}

void
simulator3SetUp::init_cijs(int isub)
{
    // synthetic models M1..M22, strains d4..d9 and zprime in local indices (simulator3/simulator3SetUp.cu:768-818)
    auto* v = M[1][isub]->as<cart_volume_regular_gpu>();
    float* mp[22];
    float* dp[6];
    for (int k = 0; k < 22; ++k)
        mp[k] = data(M[k + 1], isub);
    for (int k = 0; k < 6; ++k)
        dp[k] = data(d[k + 4], isub);
    tti::abi_check(etti_init_cijs(mp, dp, zprime[isub], v->ax1()->ibeg, v->ax1()->iend, v->ax1()->ntot,
                                  v->ax2()->ntot, v->ax3()->ntot, nullptr),
                   "init_cijs");
    // The reference fills x in [ixbeg, ixend] only, so the models' x-ghosts stay 0 and the density image, which
    // divides by M18 at stencil neighbours, is NaN next to every x-face. The models' x-ghosts take the same formula
    // here (the strains keep the reference's fill); with zero coefficients every field is unchanged by this.
    tti::abi_check(etti_init_cijs(mp, nullptr, zprime[isub], 0, v->ax1()->ntot - 1, v->ax1()->ntot, v->ax2()->ntot,
                                  v->ax3()->ntot, nullptr),
                   "init_cijs");
}

void
simulator3SetUp::set_coefficients(const float* df, const float* si)
{
    tti::abi_check(etti_init_constant(df, si), "etti_init_constant");
}

void
simulator3SetUp::initializeSource(float p)
{
    // f4..f9 = p * (1..6) at the global centre (simulator3/simulator3SetUp.cu:2659-2689)
    for (int i = 0; i < nsubs; ++i) {
        int sid[3], loc[3];
        decompmgr->getSplitLocalSubDomID(i, sid);
        bool inside = true;
        for (int a = 0; a < 3; ++a) {
            const int st = decompmgr->getOffset(sid[a], a);
            const int c = n_total[a] / 2;
            inside = inside && st <= c && c < st + decompmgr->getNumPtsSplit(sid[a], a);
            loc[a] = c - st + radius;
        }
        if (!inside)
            continue;
        auto* v = f[4][i]->as<cart_volume_regular_gpu>();
        const long idx = ((long)loc[2] * v->ax2()->ntot + loc[1]) * v->ax1()->ntot + loc[0];
        tti::abi_check(etti_init_src(data(f[4], i), data(f[5], i), data(f[6], i), data(f[7], i), data(f[8], i),
                                     data(f[9], i), idx, p, nullptr),
                       "initializeSource");
    }
}

void
simulator3SetUp::computeP(int isub)
{
    auto* p = p_gpu[isub]->as<cart_volume_regular_gpu>();
    tti::abi_check(etti_compute_p(p->getData(), data(f[4], isub), data(f[5], isub), data(f[6], isub),
                                  data(f[7], isub), data(f[8], isub), data(f[9], isub), (long)p->nelems(), nullptr),
                   "computeP");
}

void
simulator3SetUp::halo(haloManager3D_gpu* h)
{
    timer_.time(times_.halo_s, [&] {
        h->start_update();
        h->finish_update();
    });
}

namespace {
struct dims
{
    int nx, ny, nz, ldx, ldy;
};
dims
dims_of(cart_volume<float>* v)
{
    auto* g = v->as<cart_volume_regular_gpu>();
    return { g->ax1()->n, g->ax2()->n, g->ax3()->n, g->ax1()->ntot, g->ax2()->ntot };
}
} // namespace

void
simulator3SetUp::velocity_all()
{
    timer_.time(times_.velocity_s, [&] {
        for (int isub = 0; isub < nsubs; ++isub) {
            const dims D = dims_of(f[1][isub]);
            auto p = [&](int k) { return data(f[k], isub); };
            tti::abi_check(vPlainLoop(p(1), p(2), p(3), p(4), p(5), p(6), p(7), p(8), p(9), data(M[22], isub),
                                      zprime[isub], D.nx, D.ny, D.nz, D.ldx, D.ldy, nullptr, nullptr),
                           "vPlainLoop");
        }
    });
}

void
simulator3SetUp::strain_all()
{
    timer_.time(times_.strain_s, [&] {
        for (int isub = 0; isub < nsubs; ++isub) {
            const dims D = dims_of(f[1][isub]);
            auto q = [&](int k) { return data(d[k], isub); };
            tti::abi_check(vgPlainLoop(data(f[1], isub), data(f[2], isub), data(f[3], isub), q(4), q(5), q(6), q(7),
                                       q(8), q(9), zprime[isub], D.nx, D.ny, D.nz, D.ldx, D.ldy, nullptr, nullptr),
                           "vgPlainLoop");
        }
    });
}

void
simulator3SetUp::stress_all()
{
    timer_.time(times_.stress_s, [&] {
        for (int isub = 0; isub < nsubs; ++isub) {
            const dims D = dims_of(f[4][isub]);
            auto p = [&](int k) { return data(f[k], isub); };
            auto m = [&](int k) { return data(M[k], isub); };
            auto q = [&](int k) { return data(d[k], isub); };
            tti::abi_check(sPlainLoop(p(4), p(5), p(6), p(7), p(8), p(9), m(1), m(2), m(3), m(4), m(5), m(6), m(7),
                                      m(8), m(9), m(10), m(11), m(12), m(13), m(14), m(15), m(16), m(17), m(18),
                                      m(19), m(20), m(21), q(4), q(5), q(6), q(7), q(8), q(9), p(10), p(11), p(12),
                                      p(13), p(14), p(15), p(16), p(17), p(18), p(19), p(20), p(21), D.nx, D.ny,
                                      D.nz, D.ldx, D.ldy, nullptr, nullptr),
                           "sPlainLoop");
        }
    });
}

void
simulator3SetUp::mirrors()
{
    // mirror_s twice, then mirror_v twice, as the reference issues them (simulator3/simulator3SetUp.cu:2250-2264);
    // the mirrors are idempotent
    if (!free_surface_)
        return;
    for (int rep = 0; rep < 2; ++rep)
        for (int isub = 0; isub < nsubs; ++isub)
            if (top_face_[isub]) {
                auto* v = f[4][isub]->as<cart_volume_regular_gpu>();
                const long plane = (long)v->ax1()->ntot * v->ax2()->ntot;
                tti::abi_check(mirror_s_inner(data(f[4], isub), data(f[5], isub), data(f[6], isub),
                                              data(f[7], isub), data(f[8], isub), data(f[9], isub), plane,
                                              v->ax3()->ibeg, nullptr),
                               "mirror_s");
            }
    for (int rep = 0; rep < 2; ++rep)
        for (int isub = 0; isub < nsubs; ++isub)
            if (top_face_[isub]) {
                auto* v = f[1][isub]->as<cart_volume_regular_gpu>();
                const long plane = (long)v->ax1()->ntot * v->ax2()->ntot;
                tti::abi_check(mirror_v_inner(data(f[1], isub), data(f[2], isub), data(f[3], isub), plane,
                                              v->ax3()->ibeg, nullptr),
                               "mirror_v");
            }
}

void
simulator3SetUp::update_g_v(int isub)
{
    // g1 += f1 snap_f1 + f3 snap_f3 (simulator3/simulator3SetUp.cu:1502-1538)
    auto* a = f[1][isub]->as<cart_volume_regular_gpu>();
    auto* s = snap_f[1][isub]->as<cart_volume_regular_gpu>();
    tti::abi_check(update_gv_inner(s->getData(), snap_f[3][isub]->as<cart_volume_regular_gpu>()->getData(),
                                   data(g1, isub), a->getData(), data(f[3], isub), a->ax1()->ibeg, a->ax1()->iend,
                                   a->ax2()->ibeg, a->ax2()->iend, a->ax3()->ibeg, a->ax3()->iend, s->ax1()->ntot,
                                   s->ax2()->ntot, a->ax1()->ntot, a->ax2()->ntot, nullptr),
                   "update_g_v");
}

void
simulator3SetUp::update_g_r(int isub)
{
    // update_g_r_impl(isub, d4..d9, g2, f4, f6, M18, M1..M4, M5..M7, d1, d2, d3) (simulator3/simulator3SetUp.cu:
    // 2090-2128, called at :2627-2631): (d4, d5, d6) and (d7, d8, d9) are the x, y, z derivatives of the two fields
    auto* v = f[4][isub]->as<cart_volume_regular_gpu>();
    auto* g = g2[isub]->as<cart_volume_regular_gpu>();
    auto q = [&](int k) { return data(d[k], isub); };
    auto m = [&](int k) { return data(M[k], isub); };
    tti::abi_check(launch_update_g_r(g->getData(), q(4), q(7), q(5), q(8), q(6), q(9), data(f[4], isub),
                                     data(f[6], isub), m(18), m(1), m(2), m(3), m(4), m(5), m(6), m(7),
                                     v->ax1()->ibeg, v->ax1()->iend, v->ax2()->ibeg, v->ax2()->iend, v->ax3()->ibeg,
                                     v->ax3()->iend, v->ax1()->ntot, v->ax2()->ntot, g->ax1()->ntot, g->ax2()->ntot,
                                     d1, d2, d3, 8, 0, nullptr),
                   "update_g_r");
}

void
simulator3SetUp::forward(bool has_src_p)
{
    // simulator3/simulator3SetUp.cu:2190-2418
    if (forward_done_)
        throw std::runtime_error("simulator3SetUp: the forward pass runs once per object");
    forward_done_ = true;
    if (has_src_p)
        initializeSource(500.0f);
    next_snap = 0;
    for (int iter = 0; iter < _niter; ++iter) {
        halo(halomgr_s);
        velocity_all();
        halo(halomgr_v);
        mirrors();
        strain_all();
        halo(halomgr_vd);
        stress_all();
        for (int isub = 0; isub < nsubs; ++isub)
            computeP(isub);
        if (_fwd_only || iter != next_snap)
            continue;
        snap_lock_write();
        for (int isub = 0; isub < nsubs; ++isub)
            for (int k = 1; k <= 3; ++k)
                snap_f[k][isub]->as<cart_volume_regular_gpu>()->copyData(f[k][isub], true);
        halo(halomgr_s);
        for (int isub = 0; isub < nsubs; ++isub)
            for (int k = 4; k <= 9; ++k)
                snap_f[k][isub]->as<cart_volume_regular_gpu>()->copyData(f[k][isub], false);
        tti::hip_check(hipEventRecord(writeSnapshotsCompleteEvent_, nullptr), "hipEventRecord");
        snap_start_write();
        next_snap += _xcorr_step;
    }
}

void
simulator3SetUp::adjoint()
{
    // simulator3/simulator3SetUp.cu:2413-2641
    if (!forward_done_ || _fwd_only)
        throw std::runtime_error("simulator3SetUp: the adjoint pass needs the forward snapshots");
    snap_start_read();
    next_snap -= _xcorr_step;
    for (int iter = _niter - 1; iter >= 0; --iter) {
        halo(halomgr_v);
        stress_all();
        halo(halomgr_s);
        strain_all();
        halo(halomgr_vd);
        mirrors();
        velocity_all();
        if (iter != next_snap)
            continue;
        snap_finish_read();
        for (int isub = 0; isub < nsubs; ++isub)
            update_g_v(isub);
        halo(halomgr_v);
        for (int isub = 0; isub < nsubs; ++isub)
            update_g_r(isub);
        tti::hip_check(hipEventRecord(readSnapshotsCompleteEvent_, nullptr), "hipEventRecord");
        next_snap -= _xcorr_step;
        if (next_snap >= 0)
            snap_start_read();
    }
    tti::hip_check(hipDeviceSynchronize(), "simulator3SetUp::adjoint");
}

void
simulator3SetUp::execute(bool has_src_p, bool /*has_qp_*/)
{
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
        std::printf("Max time for velocity gradient: %g\n", times_.strain_s);
        std::printf("Max time for stress: %g\n", times_.stress_s);
        std::printf("Max time for halo-exchange: %g\n", times_.halo_s);
    }
}

cart_volume_regular_gpu*
simulator3SetUp::volume(const std::string& name, int isub)
{
    if (isub < 0 || isub >= nsubs)
        throw std::runtime_error("volume: sub-volume index out of range");
    auto num = [&](const char* prefix, int lo, int hi) -> int {
        const size_t n = std::char_traits<char>::length(prefix);
        if (name.compare(0, n, prefix) != 0 || name.size() == n)
            return -1;
        for (size_t i = n; i < name.size(); ++i)
            if (name[i] < '0' || name[i] > '9')
                return -1;
        const int k = std::stoi(name.substr(n));
        return k >= lo && k <= hi ? k : -1;
    };
    int k;
    if ((k = num("snap_f", 1, 9)) > 0)
        return snap_f[k][isub]->as<cart_volume_regular_gpu>();
    if ((k = num("f", 1, 21)) > 0)
        return f[k][isub]->as<cart_volume_regular_gpu>();
    if ((k = num("M", 1, 22)) > 0)
        return M[k][isub]->as<cart_volume_regular_gpu>();
    if ((k = num("d", 4, 9)) > 0)
        return d[k][isub]->as<cart_volume_regular_gpu>();
    if (name == "p")
        return p_gpu[isub]->as<cart_volume_regular_gpu>();
    if (name == "g1")
        return g1[isub]->as<cart_volume_regular_gpu>();
    if (name == "g2")
        return g2[isub]->as<cart_volume_regular_gpu>();
    throw std::runtime_error("volume: unknown name " + name);
}

void
simulator3SetUp::set_snapshot_store(const std::string& scratch, int max_mem_steps)
{
    if (forward_done_)
        throw std::runtime_error("simulator3SetUp: the snapshot store is set before the forward pass");
    long nfloats = 0;
    for (int n : corrBuff_size)
        nfloats += n;
    scratch_folder = scratch.empty() ? std::string(".") : scratch;
    file_snap_p.reset(); // the old store (and its scratch file) goes first
    file_snap_p.reset(file_snap_factory::create(_comm, const_cast<char*>(scratch_folder.c_str()), (uint64_t)nfloats,
                                                memoryPercent, numSnaps, 0, max_mem_steps));
}

void
simulator3SetUp::snap_lock_write()
{
    if (useZFP_)
        snapReaderWriter->lock_for_write_compressed_fwi_snapshot();
    else
        snapReaderWriter->lock_for_write_uncompressed_fwi_snapshot();
}

void
simulator3SetUp::snap_start_write()
{
    if (useZFP_)
        snapReaderWriter->start_write_compressed_fwi_snapshot(corrBuffList, corrBuff_size, file_snap_p.get(),
                                                              zfp_.zipped, zfp_.fields, zfp_.stream);
    else
        snapReaderWriter->start_write_uncompressed_fwi_snapshot(corrBuffList, corrBuff_size, file_snap_p.get());
}

void
simulator3SetUp::snap_start_read()
{
    if (useZFP_)
        snapReaderWriter->start_read_compressed_fwi_snapshot(corrBuffList, corrBuff_size, file_snap_p.get(),
                                                             zfp_.zipped, zfp_.fields, zfp_.stream);
    else
        snapReaderWriter->start_read_uncompressed_fwi_snapshot(corrBuffList, corrBuff_size, file_snap_p.get());
}

void
simulator3SetUp::snap_finish_read()
{
    if (useZFP_)
        snapReaderWriter->finish_read_compressed_fwi_snapshot();
    else
        snapReaderWriter->finish_read_uncompressed_fwi_snapshot();
}
