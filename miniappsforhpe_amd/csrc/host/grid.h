// Grid and memory layer of the host side: axis, volume_index and the cart_volume family, with the member and method
// names of the reference's infrastructure/ classes so driver code written against them compiles unchanged.
//
// Reference: infrastructure/axis.h:36-259, axis.cpp:19-100 (padding and ghost rules); utils/volume_index.h:7-27;
// infrastructure/cart_volume.h:44-75, cart_volume_regular.h:11-28, cart_volume_regular_gpu.h:20-61,
// cart_volume_regular_gpu.cu:11-382, cart_volume_regular_host.cu.
//
// Differences (deliberate): HIP errors throw tti::hip_error (a std::runtime_error) instead of aborting, so a host
// process can report them; memory is hipMalloc'd / hipHostMalloc'd (pinned) and always zeroed on allocation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

namespace tti {

struct hip_error : std::runtime_error
{
    hip_error(hipError_t e, const char* where)
        : std::runtime_error(std::string(where) + ": " + hipGetErrorString(e)), code(e)
    {}
    hipError_t code;
};

inline void
hip_check(hipError_t e, const char* where)
{
    if (e != hipSuccess)
        throw hip_error(e, where);
}

// status code of the C ABI -> exception
inline void
abi_check(int rc, const char* where)
{
    if (rc != 0)
        throw std::runtime_error(std::string(where) + " failed with status " + std::to_string(rc));
}

} // namespace tti

enum class AlignMemBytes
{
    NOEXTENSION = 0,
    CACHELINE = 128,
    VEC16 = 16
};

struct AlignmentElem
{
    AlignmentElem() : alignValue(AlignMemBytes::NOEXTENSION), nElems(0) {}
    AlignmentElem(AlignMemBytes alm, int sizeOfElem)
        : alignValue(alm), nElems(sizeOfElem > 0 ? static_cast<unsigned>(alm) / sizeOfElem : 0)
    {}
    AlignMemBytes alignValue;
    unsigned nElems;
};

// One axis of a volume: n interior points with origin o and step d, nghost ghosts on each side, padded with
// trailing elements so that ntot is a multiple of alignMnt.nElems.
class axis
{
public:
    axis(float o_, float d_, int n_, AlignmentElem alignMem = AlignmentElem());
    axis(float o_, float e_, float d_, AlignmentElem alignMem = AlignmentElem());
    axis(float o_, float d_, int n_, int nghost_, AlignmentElem alignMem = AlignmentElem());
    explicit axis(const axis* axin);

    float o = 0, d = 1, e = 0, opad = 0, epad = 0, invd = 1;
    int n = 0, nghost = 0, ntot = 0, npad_trailing = 0, ibeg = 0, iend = -1, nvalid = 0, nallocated_elements = 0;
    AlignmentElem alignMnt;
};

struct volume_index
{
    volume_index() = delete;
    __host__ __device__ volume_index(int _n0, int _n1 = 1, int _n2 = 1) : n0(_n0), n1(_n1), n2(_n2) {}
    int n0, n1, n2;
    template <typename T>
    __host__ __device__ T& operator()(T* data, int i0, int i1, int i2) const
    {
        return data[((int64_t)i2 * n1 + i1) * n0 + i0];
    }
};

template <class T>
class cart_volume
{
public:
    virtual ~cart_volume() = default;
    template <class V>
    V* as()
    {
        V* p = dynamic_cast<V*>(this);
        if (!p)
            throw std::runtime_error("cart_volume: wrong volume type");
        return p;
    }
};

class cart_volume_regular : public cart_volume<float>
{
public:
    cart_volume_regular(const axis* a1, const axis* a2, const axis* a3);
    axis* ax1() const { return a1_.get(); }
    axis* ax2() const { return a2_.get(); }
    axis* ax3() const { return a3_.get(); }
    size_t nelems() const { return (size_t)a1_->ntot * a2_->ntot * a3_->ntot; }

protected:
    std::unique_ptr<axis> a1_, a2_, a3_;
};

class cart_volume_regular_host;

class cart_volume_regular_gpu : public cart_volume_regular
{
public:
    using T = float;
    cart_volume_regular_gpu() = delete;
    cart_volume_regular_gpu(const cart_volume_regular_gpu&) = delete;
    cart_volume_regular_gpu& operator=(const cart_volume_regular_gpu&) = delete;
    cart_volume_regular_gpu(const axis* ax1_, const axis* ax2_, const axis* ax3_, bool set_to_zero = false);
    ~cart_volume_regular_gpu() override;

    volume_index vol_idx() const { return volume_index(a1_->ntot, a2_->ntot, a3_->ntot); }
    T* getData() { return data; }
    size_t getSize() const { return data_size; }
    bool isAligned() const;
    void set_constant(float val, bool skipHalos = true, bool skipInterleaving = false);
    void getDims(int& n0, int& n1, int& n2, bool skipHalosAndPadding = false) const;
    // this <- vol (device to device); skipHalos copies interior to interior
    void copyData(cart_volume<float>* vol, bool skipHalos = true, bool skipInterleaving = false);
    void copyFrom(cart_volume_regular_host* from, bool skipHalos = true);
    void copyTo(cart_volume_regular_host* to, bool skipHalos = true);
    void copyFrom(cart_volume_regular_gpu* from, bool skipHalos = true);
    void copyFromAsync(cart_volume_regular_host* from, hipStream_t stream, bool skipHalos = true);
    void copyToAsync(cart_volume_regular_host* to, hipStream_t stream, bool skipHalos = true);
    void copyFromAsync(cart_volume_regular_gpu* from, hipStream_t stream, bool skipHalos = true);

private:
    T* data = nullptr;
    size_t data_size = 0;
};

// Pinned host volume (the staging buffers of the snapshot helper).
class cart_volume_regular_host : public cart_volume_regular
{
public:
    using T = float;
    cart_volume_regular_host(const axis* ax1_, const axis* ax2_, const axis* ax3_, bool set_to_zero = false);
    ~cart_volume_regular_host() override;
    cart_volume_regular_host(const cart_volume_regular_host&) = delete;
    cart_volume_regular_host& operator=(const cart_volume_regular_host&) = delete;
    T* getData() { return data; }
    size_t getSize() const { return data_size; }
    volume_index vol_idx() const { return volume_index(a1_->ntot, a2_->ntot, a3_->ntot); }

private:
    T* data = nullptr;
    size_t data_size = 0;
};

namespace tti {
// Strided copy of the box [x0, x0+nx) x ... between two regular volumes (device<->device, device<->pinned host).
void copy_box_async(const cart_volume_regular* src, int sx0, int sy0, int sz0, cart_volume_regular* dst, int dx0,
                    int dy0, int dz0, int nx, int ny, int nz, hipStream_t stream);
float* raw(cart_volume<float>* v);
bool on_device(cart_volume<float>* v);
} // namespace tti
