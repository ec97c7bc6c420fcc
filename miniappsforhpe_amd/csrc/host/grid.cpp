// Grid and memory layer (see grid.h).
#include "grid.h"

#include <cmath>
#include <cstring>

#include "tti_grid.h"

namespace {

unsigned
pad_trailing(unsigned nalm, int n)
{
    // number of trailing elements that make n a multiple of nalm (infrastructure/axis.cpp:19-27)
    if (nalm == 0 || n <= 0)
        return 0;
    return nalm * ((n - 1) / nalm + 1) - n;
}

} // namespace

axis::axis(float o_, float d_, int n_, AlignmentElem alignMem) : alignMnt(alignMem)
{
    npad_trailing = (int)pad_trailing(alignMem.nElems, n_);
    n = n_;
    o = o_;
    d = d_;
    ntot = n + npad_trailing;
    opad = o;
    e = (n - 1) * d + o;
    epad = e + npad_trailing * d;
    invd = 1.0f / d;
    nghost = 0;
    ibeg = 0;
    iend = n - 1;
    nvalid = n;
    nallocated_elements = nvalid + npad_trailing;
}

axis::axis(float o_, float e_, float d_, AlignmentElem alignMem) : alignMnt(alignMem)
{
    n = (int)std::ceil((e_ - o_) / d_ - 0.5f) + 1;
    npad_trailing = (int)pad_trailing(alignMem.nElems, n);
    ntot = n + npad_trailing;
    o = o_;
    d = d_;
    opad = o;
    e = (n - 1) * d + o;
    epad = e + npad_trailing * d;
    invd = 1.0f / d;
    nghost = 0;
    ibeg = 0;
    iend = n - 1;
    nvalid = n;
    nallocated_elements = nvalid + npad_trailing;
}

axis::axis(float o_, float d_, int n_, int nghost_, AlignmentElem alignMem) : alignMnt(alignMem)
{
    o = o_;
    d = d_;
    n = n_;
    nghost = nghost_;
    ntot = n + 2 * nghost;
    npad_trailing = (int)pad_trailing(alignMem.nElems, ntot);
    ntot += npad_trailing;
    opad = o - nghost * d;
    e = (n - 1) * d + o;
    epad = e + nghost * d + npad_trailing * d;
    invd = 1.0f / d;
    ibeg = nghost;
    iend = nghost + n - 1;
    nvalid = n + 2 * nghost;
    nallocated_elements = nvalid + npad_trailing;
}

axis::axis(const axis* a) { *this = *a; }

cart_volume_regular::cart_volume_regular(const axis* a1, const axis* a2, const axis* a3)
    : a1_(new axis(a1)), a2_(new axis(a2)), a3_(new axis(a3))
{}

cart_volume_regular_gpu::cart_volume_regular_gpu(const axis* ax1_, const axis* ax2_, const axis* ax3_, bool set_to_zero)
    : cart_volume_regular(ax1_, ax2_, ax3_)
{
    (void)set_to_zero; // device memory is always zeroed: the kernels rely on zero ghosts
    data_size = nelems();
    tti::hip_check(hipMalloc((void**)&data, data_size * sizeof(float)), "cart_volume_regular_gpu: hipMalloc");
    tti::hip_check(hipMemset(data, 0, data_size * sizeof(float)), "cart_volume_regular_gpu: hipMemset");
}

cart_volume_regular_gpu::~cart_volume_regular_gpu()
{
    if (data)
        (void)hipFree(data);
}

bool
cart_volume_regular_gpu::isAligned() const
{
    return reinterpret_cast<uintptr_t>(data) % 128 == 0 && (a1_->ntot * sizeof(float)) % 128 == 0;
}

void
cart_volume_regular_gpu::set_constant(float val, bool skipHalos, bool skipInterleaving)
{
    (void)skipInterleaving;
    if (!skipHalos) {
        tti::abi_check(tti_fill(data, (long)data_size, val, nullptr), "set_constant");
        return;
    }
    tti::abi_check(tti_fill_box(data, a1_->ntot, a2_->ntot, a1_->ibeg, a2_->ibeg, a3_->ibeg, a1_->n, a2_->n, a3_->n,
                                val, nullptr),
                   "set_constant");
}

void
cart_volume_regular_gpu::getDims(int& n0, int& n1, int& n2, bool skipHalosAndPadding) const
{
    n0 = skipHalosAndPadding ? a1_->n : a1_->ntot;
    n1 = skipHalosAndPadding ? a2_->n : a2_->ntot;
    n2 = skipHalosAndPadding ? a3_->n : a3_->ntot;
}

namespace tti {

float*
raw(cart_volume<float>* v)
{
    if (auto g = dynamic_cast<cart_volume_regular_gpu*>(v))
        return g->getData();
    if (auto h = dynamic_cast<cart_volume_regular_host*>(v))
        return h->getData();
    throw std::runtime_error("unsupported cart_volume type");
}

bool
on_device(cart_volume<float>* v)
{
    return dynamic_cast<cart_volume_regular_gpu*>(v) != nullptr;
}

void
copy_box_async(const cart_volume_regular* src, int sx0, int sy0, int sz0, cart_volume_regular* dst, int dx0, int dy0,
               int dz0, int nx, int ny, int nz, hipStream_t stream)
{
    auto* s = const_cast<cart_volume_regular*>(src);
    float* sp = raw(s);
    float* dp = raw(dst);
    const bool sdev = on_device(s), ddev = on_device(dst);
    if (sdev && ddev) {
        abi_check(tti_copy_box(sp, src->ax1()->ntot, src->ax2()->ntot, sx0, sy0, sz0, dp, dst->ax1()->ntot,
                               dst->ax2()->ntot, dx0, dy0, dz0, nx, ny, nz, stream),
                  "copy_box");
        return;
    }
    hipMemcpy3DParms p{};
    p.srcPtr = make_hipPitchedPtr(sp, (size_t)src->ax1()->ntot * sizeof(float), src->ax1()->ntot, src->ax2()->ntot);
    p.dstPtr = make_hipPitchedPtr(dp, (size_t)dst->ax1()->ntot * sizeof(float), dst->ax1()->ntot, dst->ax2()->ntot);
    p.srcPos = make_hipPos((size_t)sx0 * sizeof(float), sy0, sz0);
    p.dstPos = make_hipPos((size_t)dx0 * sizeof(float), dy0, dz0);
    p.extent = make_hipExtent((size_t)nx * sizeof(float), ny, nz);
    p.kind = hipMemcpyDefault;
    hip_check(hipMemcpy3DAsync(&p, stream), "hipMemcpy3DAsync");
}

} // namespace tti

namespace {

// dst <- src: interior to interior when skipHalos, else the whole valid region (interior + ghosts); when both
// allocations have identical shapes and halos are copied the copy is one flat transfer
// (infrastructure/cart_volume_regular_gpu.cu:172-246).
void
copy_volume(cart_volume_regular* dst, cart_volume_regular* src, bool skipHalos, hipStream_t stream)
{
    const axis *sa[3] = { src->ax1(), src->ax2(), src->ax3() }, *da[3] = { dst->ax1(), dst->ax2(), dst->ax3() };
    if (skipHalos) {
        for (int k = 0; k < 3; ++k)
            if (sa[k]->n != da[k]->n)
                throw std::runtime_error("copyData: interior sizes differ");
        tti::copy_box_async(src, sa[0]->ibeg, sa[1]->ibeg, sa[2]->ibeg, dst, da[0]->ibeg, da[1]->ibeg, da[2]->ibeg,
                            sa[0]->n, sa[1]->n, sa[2]->n, stream);
        return;
    }
    bool same = true;
    for (int k = 0; k < 3; ++k)
        same = same && sa[k]->ntot == da[k]->ntot && sa[k]->nghost == da[k]->nghost && sa[k]->n == da[k]->n;
    if (same) {
        tti::hip_check(hipMemcpyAsync(tti::raw(dst), tti::raw(src), src->nelems() * sizeof(float), hipMemcpyDefault,
                                      stream),
                       "copyData: hipMemcpyAsync");
        return;
    }
    for (int k = 0; k < 3; ++k)
        if (sa[k]->nvalid != da[k]->nvalid)
            throw std::runtime_error("copyData: valid sizes differ");
    tti::copy_box_async(src, sa[0]->ibeg - sa[0]->nghost, sa[1]->ibeg - sa[1]->nghost, sa[2]->ibeg - sa[2]->nghost, dst,
                        da[0]->ibeg - da[0]->nghost, da[1]->ibeg - da[1]->nghost, da[2]->ibeg - da[2]->nghost,
                        sa[0]->nvalid, sa[1]->nvalid, sa[2]->nvalid, stream);
}

} // namespace

void
cart_volume_regular_gpu::copyData(cart_volume<float>* vol, bool skipHalos, bool skipInterleaving)
{
    (void)skipInterleaving;
    copy_volume(this, vol->as<cart_volume_regular>(), skipHalos, nullptr);
}

void
cart_volume_regular_gpu::copyFrom(cart_volume_regular_host* from, bool skipHalos)
{
    copyFromAsync(from, nullptr, skipHalos);
    tti::hip_check(hipStreamSynchronize(nullptr), "copyFrom");
}

void
cart_volume_regular_gpu::copyTo(cart_volume_regular_host* to, bool skipHalos)
{
    copyToAsync(to, nullptr, skipHalos);
    tti::hip_check(hipStreamSynchronize(nullptr), "copyTo");
}

void
cart_volume_regular_gpu::copyFrom(cart_volume_regular_gpu* from, bool skipHalos)
{
    copyFromAsync(from, nullptr, skipHalos);
    tti::hip_check(hipStreamSynchronize(nullptr), "copyFrom");
}

void
cart_volume_regular_gpu::copyFromAsync(cart_volume_regular_host* from, hipStream_t stream, bool skipHalos)
{
    copy_volume(this, from, skipHalos, stream);
}

void
cart_volume_regular_gpu::copyToAsync(cart_volume_regular_host* to, hipStream_t stream, bool skipHalos)
{
    copy_volume(to, this, skipHalos, stream);
}

void
cart_volume_regular_gpu::copyFromAsync(cart_volume_regular_gpu* from, hipStream_t stream, bool skipHalos)
{
    copy_volume(this, from, skipHalos, stream);
}

cart_volume_regular_host::cart_volume_regular_host(const axis* ax1_, const axis* ax2_, const axis* ax3_,
                                                   bool set_to_zero)
    : cart_volume_regular(ax1_, ax2_, ax3_)
{
    data_size = nelems();
    tti::hip_check(hipHostMalloc((void**)&data, data_size * sizeof(float), hipHostMallocDefault),
                   "cart_volume_regular_host: hipHostMalloc");
    if (set_to_zero)
        std::memset(data, 0, data_size * sizeof(float));
}

cart_volume_regular_host::~cart_volume_regular_host()
{
    if (data)
        (void)hipHostFree(data);
}
