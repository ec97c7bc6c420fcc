"""TEST INFRASTRUCTURE ONLY: numpy binding of the fixed-rate ZFP oracle (libzfp_oracle.so, built from zfp_oracle.c).

Only tests/ may import this module, and only as the reference that GPU results are compared against. The product package never imports it.

Fields are float32 arrays of shape (nz, ny, ldx) or views with x contiguous; streams are uint64 word arrays.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_build" / "libzfp_oracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB))
        p, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
        L.zo_compress.restype = l
        L.zo_compress.argtypes = [p, i, i, i, l, l, d, p]
        L.zo_decompress.restype = None
        L.zo_decompress.argtypes = [p, i, i, i, l, l, d, p]
        L.zo_block_bits.restype = C.c_uint
        L.zo_block_bits.argtypes = [d]
        L.zo_fwd_lift.argtypes = [p, i]
        L.zo_inv_lift.argtypes = [p, i]
        L.zo_fwd_xform.argtypes = [p]
        L.zo_inv_xform.argtypes = [p]
        L.zo_int2uint.restype = C.c_uint32
        L.zo_int2uint.argtypes = [C.c_int32]
        L.zo_uint2int.restype = C.c_int32
        L.zo_uint2int.argtypes = [C.c_uint32]
        L.zo_perm_table.argtypes = [p]
        L.zo_encode_block.argtypes = [p, C.c_uint64, C.c_uint, p]
        L.zo_decode_block.argtypes = [p, C.c_uint64, C.c_uint, p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def block_bits(rate: float) -> int:
    return int(lib().zo_block_bits(rate))


def nblocks(nx: int, ny: int, nz: int) -> int:
    return ((nx + 3) // 4) * ((ny + 3) // 4) * ((nz + 3) // 4)


def stream_words(nx: int, ny: int, nz: int, rate: float) -> int:
    return (nblocks(nx, ny, nz) * block_bits(rate) + 63) // 64


def _strides(f: np.ndarray, nx: int):
    assert f.dtype == np.float32 and f.ndim == 3 and f.strides[2] == 4
    assert f.shape[2] >= nx
    return f.strides[1] // 4, f.strides[0] // 4


def compress(f: np.ndarray, rate: float, nx: int | None = None) -> np.ndarray:
    """Codes the (nz, ny, nx) region of f (x contiguous, any row/plane strides) at `rate` bits per value."""
    nx = f.shape[2] if nx is None else nx
    sy, sz = _strides(f, nx)
    nz, ny = f.shape[0], f.shape[1]
    words = np.zeros(stream_words(nx, ny, nz, rate), dtype=np.uint64)
    n = lib().zo_compress(_ptr(f), nx, ny, nz, sy, sz, rate, _ptr(words))
    assert n == words.size
    return words


def decompress(words: np.ndarray, shape, rate: float, out: np.ndarray | None = None) -> np.ndarray:
    nz, ny, nx = shape
    if out is None:
        out = np.zeros((nz, ny, nx), dtype=np.float32)
    sy, sz = _strides(out, nx)
    assert words.dtype == np.uint64 and words.size >= stream_words(nx, ny, nz, rate)
    lib().zo_decompress(_ptr(np.ascontiguousarray(words)), nx, ny, nz, sy, sz, rate, _ptr(out))
    return out


def fwd_lift(v) -> np.ndarray:
    a = np.array(v, dtype=np.int32)
    lib().zo_fwd_lift(_ptr(a), 1)
    return a


def inv_lift(v) -> np.ndarray:
    a = np.array(v, dtype=np.int32)
    lib().zo_inv_lift(_ptr(a), 1)
    return a


def fwd_xform(block) -> np.ndarray:
    a = np.array(block, dtype=np.int32).reshape(64).copy()
    lib().zo_fwd_xform(_ptr(a))
    return a


def inv_xform(block) -> np.ndarray:
    a = np.array(block, dtype=np.int32).reshape(64).copy()
    lib().zo_inv_xform(_ptr(a))
    return a


def int2uint(x: int) -> int:
    return int(lib().zo_int2uint(int(x)))


def uint2int(x: int) -> int:
    return int(lib().zo_uint2int(int(x)))


def perm() -> np.ndarray:
    a = np.zeros(64, dtype=np.uint8)
    lib().zo_perm_table(_ptr(a))
    return a


def encode_block(values, rate: float) -> np.ndarray:
    """One 4x4x4 block (index x + 4y + 16z) coded alone: its maxbits bits in whole words."""
    v = np.ascontiguousarray(np.asarray(values, dtype=np.float32).reshape(64))
    bits = block_bits(rate)
    words = np.zeros((bits + 63) // 64, dtype=np.uint64)
    lib().zo_encode_block(_ptr(words), 0, bits, _ptr(v))
    return words
