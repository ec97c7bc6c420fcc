// C ABI over the host classes (declared in include/tti_host.h). Exceptions become TTI_EHOST with a per-thread message.
#include "tti_host.h"

#include <cstring>
#include <string>
#include <vector>

#include "simulator2.h"
#include "simulator3.h"

struct tti_sim2
{
    simulator2SetUp* sim;
};

struct tti_sim3
{
    simulator3SetUp* sim;
};

struct tti_snap_store
{
    file_snap* store;
    long nval;
};

namespace {

thread_local std::string last_error;

template <class F>
int
guarded(F&& f)
{
    try {
        last_error.clear();
        f();
        return 0;
    } catch (const std::exception& e) {
        last_error = e.what();
    } catch (...) {
        last_error = "unknown exception";
    }
    return TTI_EHOST;
}

} // namespace

extern "C" {

const char*
tti_host_last_error(void)
{
    return last_error.c_str();
}

tti_sim2*
tti_sim2_create(int npoints, int npml, int niter, int xcorr_step, const int nzones[3], int fwd_only)
{
    return tti_sim2_create_ex(npoints, npml, niter, xcorr_step, nzones, fwd_only, 32.0f);
}

tti_sim2*
tti_sim2_create_ex(int npoints, int npml, int niter, int xcorr_step, const int nzones[3], int fwd_only,
                   float bitPerFloat)
{
    tti_sim2* h = nullptr;
    guarded([&] {
        if (!nzones)
            throw std::runtime_error("nzones is NULL");
        int nz[3] = { nzones[0], nzones[1], nzones[2] };
        int ranks[3] = { 1, 1, 1 };
        auto* s = new simulator2SetUp(npoints, npml, niter, xcorr_step, nz, ranks, 4, 4, bitPerFloat, fwd_only,
                                      MPI_COMM_WORLD);
        h = new tti_sim2{ s };
    });
    return h;
}

void
tti_sim2_destroy(tti_sim2* sim)
{
    if (!sim)
        return;
    delete sim->sim;
    delete sim;
}

int
tti_sim2_set_medium(tti_sim2* sim, float vp, float eps, float delta, float vs, float theta, float phi, float rho)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->set_medium(vp, eps, delta, vs, theta, phi, rho); });
}

int
tti_sim2_set_snapshot_store(tti_sim2* sim, const char* scratch, int max_mem_steps)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->set_snapshot_store(scratch ? scratch : ".", max_mem_steps); });
}

int
tti_sim2_set_free_surface(tti_sim2* sim, int on)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->set_free_surface(on != 0); });
}

int
tti_sim2_set_fused(tti_sim2* sim, int on)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->set_fused(on != 0); });
}

int
tti_sim2_set_update_born(tti_sim2* sim, int on)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->set_update_born(on != 0); });
}

int
tti_sim2_fused_active(tti_sim2* sim)
{
    if (!sim)
        return -1;
    int on = 0;
    const int rc = guarded([&] { on = sim->sim->fused_active() ? 1 : 0; });
    return rc != 0 ? rc : on;
}

int
tti_sim2_execute(tti_sim2* sim, int has_src_p, int has_qp)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->execute(has_src_p != 0, has_qp != 0); });
}

int
tti_sim2_forward(tti_sim2* sim, int has_src_p)
{
    if (!sim)
        return -1;
    return guarded([&] {
        sim->sim->forward(has_src_p != 0);
        tti::hip_check(hipDeviceSynchronize(), "forward");
    });
}

int
tti_sim2_adjoint(tti_sim2* sim)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->adjoint(); });
}

int
tti_sim2_num_subs(const tti_sim2* sim)
{
    return sim ? sim->sim->num_subs() : -1;
}

float
tti_sim2_dt(const tti_sim2* sim)
{
    return sim ? sim->sim->get_dt() : 0.0f;
}

int
tti_sim2_volume_info(tti_sim2* sim, const char* name, int isub, int info[12])
{
    if (!sim || !name || !info)
        return -1;
    return guarded([&] {
        auto* v = sim->sim->volume(name, isub);
        const auto* dm = sim->sim->decomposition();
        int sid[3];
        dm->getSplitLocalSubDomID(isub, sid);
        const axis* ax[3] = { v->ax1(), v->ax2(), v->ax3() };
        for (int d = 0; d < 3; ++d) {
            info[d] = dm->getOffset(sid[d], d);
            info[3 + d] = ax[d]->n;
            info[6 + d] = ax[d]->ntot;
            info[9 + d] = ax[d]->ibeg;
        }
    });
}

int
tti_sim2_read(tti_sim2* sim, const char* name, int isub, float* host_out)
{
    if (!sim || !name || !host_out)
        return -1;
    return guarded([&] {
        auto* v = sim->sim->volume(name, isub);
        tti::hip_check(hipDeviceSynchronize(), "read");
        tti::hip_check(hipMemcpy(host_out, v->getData(), v->nelems() * sizeof(float), hipMemcpyDeviceToHost),
                       "read: hipMemcpy");
    });
}

int
tti_sim2_write(tti_sim2* sim, const char* name, int isub, const float* host_in, long nfloats)
{
    if (!sim || !name || !host_in)
        return -1;
    return guarded([&] {
        auto* v = sim->sim->volume(name, isub);
        if (nfloats != (long)v->nelems())
            throw std::runtime_error(std::string("write ") + name + ": " + std::to_string(nfloats) +
                                     " floats for a volume of " + std::to_string(v->nelems()));
        tti::hip_check(hipDeviceSynchronize(), "write");
        tti::hip_check(hipMemcpy(v->getData(), host_in, v->nelems() * sizeof(float), hipMemcpyHostToDevice),
                       "write: hipMemcpy");
    });
}

int
tti_sim2_timings(const tti_sim2* sim, double out[5])
{
    if (!sim || !out)
        return -1;
    const auto& t = sim->sim->get_timings();
    out[0] = t.forward_s;
    out[1] = t.adjoint_s;
    out[2] = t.velocity_s;
    out[3] = t.stress_s;
    out[4] = t.halo_s;
    return 0;
}

// ---- simulator3SetUp ---------------------------------------------------------------------------------------------

tti_sim3*
tti_sim3_create(int npoints, int npml, int niter, int xcorr_step, const int nzones[3], int fwd_only)
{
    return tti_sim3_create_ex(npoints, npml, niter, xcorr_step, nzones, fwd_only, 32.0f);
}

tti_sim3*
tti_sim3_create_ex(int npoints, int npml, int niter, int xcorr_step, const int nzones[3], int fwd_only,
                   float bitPerFloat)
{
    tti_sim3* h = nullptr;
    guarded([&] {
        if (!nzones)
            throw std::runtime_error("nzones is NULL");
        int nz[3] = { nzones[0], nzones[1], nzones[2] };
        int ranks[3] = { 1, 1, 1 };
        auto* s = new simulator3SetUp(npoints, npml, niter, xcorr_step, nz, ranks, 4, 4, bitPerFloat, fwd_only,
                                      MPI_COMM_WORLD);
        h = new tti_sim3{ s };
    });
    return h;
}

void
tti_sim3_destroy(tti_sim3* sim)
{
    if (!sim)
        return;
    delete sim->sim;
    delete sim;
}

int
tti_sim3_set_coefficients(tti_sim3* sim, const float* df, const float* si)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->set_coefficients(df, si); });
}

int
tti_sim3_set_snapshot_store(tti_sim3* sim, const char* scratch, int max_mem_steps)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->set_snapshot_store(scratch ? scratch : ".", max_mem_steps); });
}

int
tti_sim3_set_free_surface(tti_sim3* sim, int on)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->set_free_surface(on != 0); });
}

int
tti_sim3_execute(tti_sim3* sim, int has_src_p, int has_qp)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->execute(has_src_p != 0, has_qp != 0); });
}

int
tti_sim3_forward(tti_sim3* sim, int has_src_p)
{
    if (!sim)
        return -1;
    return guarded([&] {
        sim->sim->forward(has_src_p != 0);
        tti::hip_check(hipDeviceSynchronize(), "forward");
    });
}

int
tti_sim3_adjoint(tti_sim3* sim)
{
    if (!sim)
        return -1;
    return guarded([&] { sim->sim->adjoint(); });
}

int
tti_sim3_num_subs(const tti_sim3* sim)
{
    return sim ? sim->sim->num_subs() : -1;
}

int
tti_sim3_volume_info(tti_sim3* sim, const char* name, int isub, int info[12])
{
    if (!sim || !name || !info)
        return -1;
    return guarded([&] {
        auto* v = sim->sim->volume(name, isub);
        const auto* dm = sim->sim->decomposition();
        int sid[3];
        dm->getSplitLocalSubDomID(isub, sid);
        const axis* ax[3] = { v->ax1(), v->ax2(), v->ax3() };
        for (int d = 0; d < 3; ++d) {
            info[d] = dm->getOffset(sid[d], d);
            info[3 + d] = ax[d]->n;
            info[6 + d] = ax[d]->ntot;
            info[9 + d] = ax[d]->ibeg;
        }
    });
}

int
tti_sim3_read(tti_sim3* sim, const char* name, int isub, float* host_out)
{
    if (!sim || !name || !host_out)
        return -1;
    return guarded([&] {
        tti::hip_check(hipDeviceSynchronize(), "read");
        if (std::string(name) == "zprime") {
            const int ldz = sim->sim->volume("f1", isub)->ax3()->ntot;
            tti::hip_check(hipMemcpy(host_out, sim->sim->zprime_data(isub), ldz * sizeof(float),
                                     hipMemcpyDeviceToHost),
                           "read: hipMemcpy");
            return;
        }
        auto* v = sim->sim->volume(name, isub);
        tti::hip_check(hipMemcpy(host_out, v->getData(), v->nelems() * sizeof(float), hipMemcpyDeviceToHost),
                       "read: hipMemcpy");
    });
}

int
tti_sim3_timings(const tti_sim3* sim, double out[6])
{
    if (!sim || !out)
        return -1;
    const auto& t = sim->sim->get_timings();
    out[0] = t.forward_s;
    out[1] = t.adjoint_s;
    out[2] = t.velocity_s;
    out[3] = t.strain_s;
    out[4] = t.stress_s;
    out[5] = t.halo_s;
    return 0;
}

tti_snap_store*
tti_snap_store_create(long nval, float mem_perc, int nt)
{
    tti_snap_store* h = nullptr;
    guarded([&] {
        char scratch[] = ".";
        h = new tti_snap_store{ file_snap_factory::create(MPI_COMM_WORLD, scratch, (uint64_t)nval, mem_perc, nt, 0),
                                nval };
    });
    return h;
}

tti_snap_store*
tti_snap_store_create_ex(long nval, float mem_perc, int nt, const char* scratch, int max_mem_steps)
{
    tti_snap_store* h = nullptr;
    guarded([&] {
        std::string dir = scratch && *scratch ? scratch : ".";
        h = new tti_snap_store{ file_snap_factory::create(MPI_COMM_WORLD, &dir[0], (uint64_t)nval, mem_perc, nt, 0,
                                                          max_mem_steps),
                                nval };
    });
    return h;
}

int
tti_snap_store_strategy(const tti_snap_store* s, char* out, int len, int* direct_io)
{
    if (!s || !out || len <= 0)
        return -1;
    return guarded([&] {
        auto* o = dynamic_cast<const file_snap_orig*>(s->store);
        if (!o)
            throw std::runtime_error("tti_snap_store_strategy: not a file_snap_orig store");
        const std::string st = o->strategy_string();
        if ((int)st.size() >= len)
            throw std::runtime_error("tti_snap_store_strategy: output buffer too small");
        std::memcpy(out, st.c_str(), st.size() + 1);
        if (direct_io)
            *direct_io = o->direct_io() ? 1 : 0;
    });
}

int
tti_snap_store_write(tti_snap_store* s, const float* step)
{
    if (!s || !step)
        return -1;
    return guarded([&] { s->store->write_step(const_cast<float*>(step)); });
}

int
tti_snap_store_read(tti_snap_store* s, float* step)
{
    if (!s || !step)
        return -1;
    return guarded([&] { s->store->read_step(step); });
}

void
tti_snap_store_destroy(tti_snap_store* s)
{
    if (!s)
        return;
    delete s->store;
    delete s;
}

} // extern "C"
