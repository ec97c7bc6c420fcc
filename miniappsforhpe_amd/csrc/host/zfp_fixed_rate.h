// The part of the zfp C API (LLNL) that the reference's compressed-snapshot path calls, over the gfx950 fixed-rate
// codec of tti_zfp.h: 3-D float fields, fixed-rate mode, GPU execution.
//
// Call sites this serves: simulator2/simulator2SetUp.cu:285-323 (zfp_stream_open, zfp_stream_set_rate,
// zfp_field_3d, zfp_field_set_stride_3d, zfp_stream_maximum_size, zfp_stream_set_execution),
// snap_helpers/rtm_snap_gpu.cu:228-280 (zfp_field_set_pointer, stream_open, zfp_stream_set_bit_stream,
// zfp_stream_rewind, zfp_compress, zfp_decompress, stream_close) and simulator3/simulator3SetUp.cu:583-620.
//
// Everything here is inline, so the library exports no zfp_* symbol and cannot collide with a real libzfp in the
// same process. Semantics kept: the field pointer is device memory (exec policy "cuda", here the HIP device); the
// bit stream buffer may be pinned host or device memory; zfp_compress / zfp_decompress return once the data are in
// place (0 on failure, as the library does). Not provided: other types, dimensions and modes, headers, serial/OpenMP
// execution. Work runs on the stream given by zfp_stream_set_hip_stream (default: the null stream, as the reference's
// zfpCudaStream = 0).
#pragma once

#include <cstddef>
#include <hip/hip_runtime.h>

#include "tti_zfp.h"

enum zfp_type
{
    zfp_type_none = 0,
    zfp_type_int32 = 1,
    zfp_type_int64 = 2,
    zfp_type_float = 3,
    zfp_type_double = 4
};

enum zfp_exec_policy
{
    zfp_exec_serial = 0,
    zfp_exec_omp = 1,
    zfp_exec_cuda = 2 // the GPU of this build
};

struct bitstream
{
    void* buffer;
    size_t bytes;
};

struct zfp_field
{
    zfp_type type;
    unsigned nx, ny, nz;
    ptrdiff_t sx, sy, sz;
    void* data;
};

struct zfp_stream
{
    bitstream* stream = nullptr;
    double rate = 0.0; // 0: not in fixed-rate mode
    zfp_exec_policy exec = zfp_exec_serial;
    hipStream_t hip_stream = nullptr;
    void* staging = nullptr; // device copy of a host bit stream, grown on demand
    size_t staging_bytes = 0;
};

inline bitstream*
stream_open(void* buffer, size_t bytes)
{
    return new bitstream{ buffer, bytes };
}

inline void
stream_close(bitstream* s)
{
    delete s;
}

inline zfp_stream*
zfp_stream_open(bitstream* s)
{
    auto* z = new zfp_stream;
    z->stream = s;
    return z;
}

inline void
zfp_stream_close(zfp_stream* z)
{
    if (!z)
        return;
    if (z->staging)
        (void)hipFree(z->staging);
    delete z;
}

// Fixed-rate mode; returns the rate actually used (bits per block / 64). Only 3-D float fields are coded, and
// word-aligned blocks (wra != 0) are not provided.
inline double
zfp_stream_set_rate(zfp_stream* z, double rate, zfp_type type, unsigned dims, int wra)
{
    if (!z || type != zfp_type_float || dims != 3 || wra || !(rate > 0.0))
        return 0.0;
    z->rate = rate;
    return tti_zfp_block_bits(rate) / 64.0;
}

inline int
zfp_stream_set_execution(zfp_stream* z, zfp_exec_policy policy)
{
    if (!z || policy != zfp_exec_cuda)
        return 0;
    z->exec = policy;
    return 1;
}

// Not in zfp: the HIP stream the codec runs on (the library's CUDA back end uses the default stream).
inline void
zfp_stream_set_hip_stream(zfp_stream* z, hipStream_t s)
{
    z->hip_stream = s;
}

inline void
zfp_stream_set_bit_stream(zfp_stream* z, bitstream* s)
{
    z->stream = s;
}

inline void
zfp_stream_rewind(zfp_stream*)
{
    // every call codes from the start of the buffer
}

inline zfp_field*
zfp_field_3d(void* p, zfp_type type, unsigned nx, unsigned ny, unsigned nz)
{
    return new zfp_field{ type, nx, ny, nz, 0, 0, 0, p };
}

inline void
zfp_field_set_stride_3d(zfp_field* f, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz)
{
    f->sx = sx;
    f->sy = sy;
    f->sz = sz;
}

inline void
zfp_field_set_pointer(zfp_field* f, void* p)
{
    f->data = p;
}

inline void
zfp_field_free(zfp_field* f)
{
    delete f;
}

inline size_t
zfp_stream_maximum_size(const zfp_stream* z, const zfp_field* f)
{
    if (!z || !f || z->rate <= 0.0 || f->type != zfp_type_float)
        return 0;
    return tti_zfp_maximum_size(z->rate, (int)f->nx, (int)f->ny, (int)f->nz);
}

namespace tti_zfp_detail {

// row and plane strides in elements (zero strides mean contiguous, as in zfp); x must be contiguous
inline bool
strides(const zfp_field* f, long& sy, long& sz)
{
    if (f->sx != 0 && f->sx != 1)
        return false;
    sy = f->sy ? (long)f->sy : (long)f->nx;
    sz = f->sz ? (long)f->sz : (long)f->nx * (long)f->ny;
    return true;
}

// the device buffer the codec reads or writes: the bit stream itself when it is device memory, else the staging
inline void*
device_buffer(zfp_stream* z, size_t bytes, bool& staged)
{
    hipPointerAttribute_t a;
    staged = true;
    if (hipPointerGetAttributes(&a, z->stream->buffer) == hipSuccess && a.type == hipMemoryTypeDevice)
        staged = false;
    else
        (void)hipGetLastError(); // plain host memory makes the query fail; clear it
    if (!staged)
        return z->stream->buffer;
    if (z->staging_bytes < bytes) {
        if (z->staging)
            (void)hipFree(z->staging);
        z->staging = nullptr;
        z->staging_bytes = 0;
        if (hipMalloc(&z->staging, bytes) != hipSuccess)
            return nullptr;
        z->staging_bytes = bytes;
    }
    return z->staging;
}

} // namespace tti_zfp_detail

// Codes the field into the bit stream; returns the stream size in bytes, 0 on failure.
inline size_t
zfp_compress(zfp_stream* z, const zfp_field* f)
{
    long sy, sz;
    if (!z || !f || !z->stream || z->rate <= 0.0 || f->type != zfp_type_float || !tti_zfp_detail::strides(f, sy, sz))
        return 0;
    const size_t need = tti_zfp_stream_bytes(z->rate, (int)f->nx, (int)f->ny, (int)f->nz);
    if (need == 0 || z->stream->bytes < need)
        return 0;
    bool staged;
    void* dev = tti_zfp_detail::device_buffer(z, need, staged);
    if (!dev)
        return 0;
    size_t out = 0;
    if (tti_zfp_compress_3d(static_cast<const float*>(f->data), (int)f->nx, (int)f->ny, (int)f->nz, sy, sz, z->rate,
                            dev, need, &out, z->hip_stream) != 0)
        return 0;
    if (staged && hipMemcpyAsync(z->stream->buffer, dev, out, hipMemcpyDeviceToHost, z->hip_stream) != hipSuccess)
        return 0;
    return hipStreamSynchronize(z->hip_stream) == hipSuccess ? out : 0;
}

// Decodes the bit stream into the field; returns the bytes consumed, 0 on failure.
inline size_t
zfp_decompress(zfp_stream* z, zfp_field* f)
{
    long sy, sz;
    if (!z || !f || !z->stream || z->rate <= 0.0 || f->type != zfp_type_float || !tti_zfp_detail::strides(f, sy, sz))
        return 0;
    const size_t need = tti_zfp_stream_bytes(z->rate, (int)f->nx, (int)f->ny, (int)f->nz);
    if (need == 0 || z->stream->bytes < need)
        return 0;
    bool staged;
    void* dev = tti_zfp_detail::device_buffer(z, need, staged);
    if (!dev)
        return 0;
    if (staged && hipMemcpyAsync(dev, z->stream->buffer, need, hipMemcpyHostToDevice, z->hip_stream) != hipSuccess)
        return 0;
    if (tti_zfp_decompress_3d(dev, need, static_cast<float*>(f->data), (int)f->nx, (int)f->ny, (int)f->nz, sy, sz,
                              z->rate, z->hip_stream) != 0)
        return 0;
    return hipStreamSynchronize(z->hip_stream) == hipSuccess ? need : 0;
}
