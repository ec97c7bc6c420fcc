/* C ABI of the fixed-rate ZFP snapshot codec, gfx950 build.
 *
 * The reference compresses each snapshot volume with the zfp library (LLNL) on the GPU: fixed-rate mode,
 * zfp_stream_set_rate(stream, bitPerFloat, zfp_type_float, 3, 0), a 3-D float field over the valid extents with
 * the volume's row and plane strides, zfp_compress / zfp_decompress (simulator2/simulator2SetUp.cu:285-323,
 * snap_helpers/rtm_snap_gpu.cu:228-280, simulator3/simulator3SetUp.cu:583-620). These entry points replace those
 * library calls for that use: the stream is the sequence of 4x4x4 blocks in raster order (x fastest), each exactly
 * tti_zfp_block_bits(rate) bits, packed LSB first into 64-bit words, as that library lays out a fixed-rate stream
 * (no header is written, as the reference writes none).
 *
 * Fields: FP32 device pointers at the first value, x contiguous, element strides sy (row) and sz (plane).
 * Streams: device pointers, 8-byte aligned. Every entry returns 0 on success, -1 for bad arguments, -1000 - hipError_t
 * for a HIP failure; `stream` is a hipStream_t (NULL = default). Calls are asynchronous on that stream.
 * Values are expected finite (NaN and infinity are not coded).
 */
#ifndef TTI_ZFP_H
#define TTI_ZFP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bits per 4x4x4 block at `rate` bits per value: floor(64 rate + 0.5), at least 9 (zfp_stream_set_rate in
 * fixed-rate mode without word alignment, as called at simulator2/simulator2SetUp.cu:292). */
unsigned tti_zfp_block_bits(double rate);

/* Bytes of the coded stream of an nx x ny x nz field: the blocks' bits rounded up to whole 64-bit words. */
size_t tti_zfp_stream_bytes(double rate, int nx, int ny, int nz);

/* Upper bound the reference allocates per volume (zfp_stream_maximum_size, simulator2/simulator2SetUp.cu:313): the
 * stream bytes plus room for a 148-bit header, rounded to whole words. */
size_t tti_zfp_maximum_size(double rate, int nx, int ny, int nz);

/* Codes the field into `out` (out_bytes >= tti_zfp_stream_bytes); *compressed_bytes (may be NULL) receives the
 * stream size. Replaces zfp_compress (snap_helpers/rtm_snap_gpu.cu:244). */
int tti_zfp_compress_3d(const float* field, int nx, int ny, int nz, long sy, long sz, double rate, void* out,
                        size_t out_bytes, size_t* compressed_bytes, void* stream);

/* Decodes a stream made by tti_zfp_compress_3d with the same sizes and rate into the field's valid region (values
 * outside it are not written). Replaces zfp_decompress (snap_helpers/rtm_snap_gpu.cu:273). */
int tti_zfp_decompress_3d(const void* in, size_t in_bytes, float* field, int nx, int ny, int nz, long sy, long sz,
                          double rate, void* stream);

#ifdef __cplusplus
}
#endif

#endif
