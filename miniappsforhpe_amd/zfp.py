"""Fixed-rate ZFP coding of 3-D FP32 device volumes (the compressed-snapshot codec, include/tti_zfp.h).

    words = compress(vol, rate)                 # vol: float32 CUDA tensor (nz, ny, ldx), region (nz, ny, nx)
    decompress(words, out, rate, nx=...)        # decodes into out's region

The stream is the fixed-rate layout of the zfp library (4x4x4 blocks in raster order, floor(64 rate + 0.5) bits
each, LSB first in 64-bit words), as the reference's snapshot store writes it (snap_helpers/rtm_snap_gpu.cu:228-280).
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import call, lib


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def block_bits(rate: float) -> int:
    return int(lib().tti_zfp_block_bits(float(rate)))


def stream_bytes(rate: float, nx: int, ny: int, nz: int) -> int:
    return int(lib().tti_zfp_stream_bytes(float(rate), nx, ny, nz))


def maximum_size(rate: float, nx: int, ny: int, nz: int) -> int:
    return int(lib().tti_zfp_maximum_size(float(rate), nx, ny, nz))


def _geometry(vol: torch.Tensor, nx: int | None):
    if vol.dtype != torch.float32 or vol.dim() != 3 or not vol.is_cuda or vol.stride(2) != 1:
        raise ValueError("expected a float32 CUDA volume (nz, ny, ldx) with x contiguous")
    nx = vol.shape[2] if nx is None else nx
    if not 0 < nx <= vol.shape[2]:
        raise ValueError("nx outside the volume")
    return nx, vol.shape[1], vol.shape[0], vol.stride(1), vol.stride(0)


def compress(vol: torch.Tensor, rate: float, nx: int | None = None, out: torch.Tensor | None = None) -> torch.Tensor:
    """Codes the (nz, ny, nx) region of vol; returns the uint64 words (an int64 tensor of the same bits)."""
    nx, ny, nz, sy, sz = _geometry(vol, nx)
    nbytes = stream_bytes(rate, nx, ny, nz)
    if out is None:
        out = torch.empty(nbytes // 8, dtype=torch.int64, device=vol.device)
    if out.dtype != torch.int64 or out.numel() * 8 < nbytes or not out.is_contiguous():
        raise ValueError("output words too small")
    used = C.c_size_t(0)
    call("tti_zfp_compress_3d", vol.data_ptr(), nx, ny, nz, sy, sz, float(rate), out.data_ptr(), out.numel() * 8,
         C.byref(used), _stream())
    return out


def decompress(words: torch.Tensor, vol: torch.Tensor, rate: float, nx: int | None = None) -> torch.Tensor:
    """Decodes words (from compress with the same region and rate) into vol's (nz, ny, nx) region."""
    nx, ny, nz, sy, sz = _geometry(vol, nx)
    if words.dtype != torch.int64 or not words.is_contiguous() or not words.is_cuda:
        raise ValueError("words must be a contiguous int64 CUDA tensor")
    call("tti_zfp_decompress_3d", words.data_ptr(), words.numel() * 8, vol.data_ptr(), nx, ny, nz, sy, sz,
         float(rate), _stream())
    return vol
