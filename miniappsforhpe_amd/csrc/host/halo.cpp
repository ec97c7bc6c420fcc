#include "halo.h"

#include <stdexcept>

namespace {

// Slab of `width` points along axis `dim` starting at array index `start`; along the other axes the interior, widened by
// the ghosts for the axes exchanged before this one (x face: interior y, z; y face: x with ghosts, interior z; z face:
// x and y with ghosts), as copy_2way_X/Y/Z do (infrastructure/haloManager3D_gpu_kernels.cuh:14-118).
struct slab
{
    int b[3], n[3];
};

slab
make_slab(const cart_volume_regular* v, int dim, int start, int width)
{
    const axis* ax[3] = { v->ax1(), v->ax2(), v->ax3() };
    slab s{};
    for (int d = 0; d < 3; ++d) {
        if (d == dim) {
            s.b[d] = start;
            s.n[d] = width;
        } else if (d < dim) {
            s.b[d] = ax[d]->ibeg - ax[d]->nghost;
            s.n[d] = ax[d]->n + 2 * ax[d]->nghost;
        } else {
            s.b[d] = ax[d]->ibeg;
            s.n[d] = ax[d]->n;
        }
    }
    return s;
}

tti_box_copy
make_copy(cart_volume_regular* src, const slab& ss, cart_volume_regular* dst, const slab& ds)
{
    for (int d = 0; d < 3; ++d)
        if (ss.n[d] != ds.n[d])
            throw std::runtime_error("haloManager3D_gpu: adjacent sub-volumes disagree on a face size");
    tti_box_copy c{};
    c.src = tti::raw(src);
    c.dst = tti::raw(dst);
    c.snelem = (long)src->nelems();
    c.dnelem = (long)dst->nelems();
    c.sldx = src->ax1()->ntot;
    c.sldy = src->ax2()->ntot;
    c.dldx = dst->ax1()->ntot;
    c.dldy = dst->ax2()->ntot;
    c.sx0 = ss.b[0];
    c.sy0 = ss.b[1];
    c.sz0 = ss.b[2];
    c.dx0 = ds.b[0];
    c.dy0 = ds.b[1];
    c.dz0 = ds.b[2];
    c.nx = ss.n[0];
    c.ny = ss.n[1];
    c.nz = ss.n[2];
    return c;
}

const axis*
axis_of(const cart_volume_regular* v, int d)
{
    return d == 0 ? v->ax1() : d == 1 ? v->ax2() : v->ax3();
}

} // namespace

haloManager3D_gpu::haloManager3D_gpu(decompositionManager3D* decomp, std::vector<cart_volume<float>**>& data,
                                     std::vector<std::vector<int>> wSz[3][2], std::vector<std::vector<bool>>&,
                                     std::vector<std::vector<bool>>&)
{
    const int nsub = decomp->getTotalNumLocalSubDom();
    const int nfields = (int)data.size();
    constexpr int neg_slot[3] = { _NH_ID_3D_NEGX_CENY_CENZ, _NH_ID_3D_CENX_NEGY_CENZ, _NH_ID_3D_CENX_CENY_NEGZ };
    for (int a = 0; a < 3; ++a)
        for (int s = 0; s < 2; ++s)
            if ((int)wSz[a][s].size() != nfields)
                throw std::runtime_error("haloManager3D_gpu: one width list per field expected");
    try {
        for (int a = 0; a < 3; ++a) {
            std::vector<tti_box_copy> copies;
            for (int j = 0; j < nsub; ++j) {
                int nh[_NH_ID_3D_TOTAL];
                decomp->getLocalNeighborID(j, nh);
                const int nbr = nh[neg_slot[a]];
                if (nbr < 0)
                    continue;
                for (int i = 0; i < nfields; ++i) {
                    auto* lo = data[i][nbr]->as<cart_volume_regular>();
                    auto* hi = data[i][j]->as<cart_volume_regular>();
                    const int r_hi = wSz[a][0][i][j];   // low ghost of the upper volume
                    const int r_lo = wSz[a][1][i][nbr]; // high ghost of the lower volume
                    const axis* la = axis_of(lo, a);
                    const axis* ha = axis_of(hi, a);
                    if (r_hi < 0 || r_lo < 0 || r_hi > ha->nghost || r_lo > la->nghost || r_hi > la->n ||
                        r_lo > ha->n)
                        throw std::runtime_error("haloManager3D_gpu: halo width exceeds a ghost zone or an interior");
                    // lower volume's last interior planes -> upper volume's low ghost
                    if (r_hi > 0)
                        copies.push_back(make_copy(lo, make_slab(lo, a, la->iend - r_hi + 1, r_hi), hi,
                                                   make_slab(hi, a, ha->ibeg - r_hi, r_hi)));
                    // upper volume's first interior planes -> lower volume's high ghost
                    if (r_lo > 0)
                        copies.push_back(make_copy(hi, make_slab(hi, a, ha->ibeg, r_lo), lo,
                                                   make_slab(lo, a, la->iend + 1, r_lo)));
                }
            }
            counts_[a] = (int)copies.size();
            plans_[a] = tti_copy_plan_create(copies.data(), (int)copies.size());
            if (!plans_[a])
                throw std::runtime_error("haloManager3D_gpu: invalid copy plan");
        }
    } catch (...) {
        for (auto*& p : plans_) {
            tti_copy_plan_destroy(p);
            p = nullptr;
        }
        throw;
    }
}

haloManager3D_gpu::~haloManager3D_gpu()
{
    for (auto* p : plans_)
        tti_copy_plan_destroy(p);
}

void
haloManager3D_gpu::start_update(hipStream_t stream)
{
    // X, then Y, then Z on one stream: each axis sees the ghosts the previous axes filled
    for (auto* p : plans_)
        tti::abi_check(tti_copy_plan_run(p, stream), "haloManager3D_gpu::start_update");
}

void
haloManager3D_gpu::finish_update(hipStream_t)
{}
