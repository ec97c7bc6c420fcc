"""The fixed-rate ZFP oracle (oracle/zfp_oracle.c) pinned by what the published algorithm fixes.

* lifting: on inputs that no right shift rounds (multiples of 2^12) the forward steps equal the published transform
  1/16 [4 4 4 4; 5 1 -1 -5; -4 4 4 -4; -2 6 -6 2] along each axis of the block, and the inverse steps undo them; a
  swapped step or sign fails the matrix identity;
* negabinary: every code is the base -2 expansion of its integer (sum of bit_i (-2)^i), and the maps invert;
* sequency order: a permutation of the 64 coefficients, non-decreasing in (i+j+k, i^2+j^2+k^2);
* known answers derived by hand from the coding rules: constant blocks of +1 and -1 at rate 8 (one significant
  coefficient 2^29 in negabinary 0x60000000 / 0x20000000, header 2*128+1 in 9 bits, then one or two significance
  runs and a zero bit pair per lower plane), and an all-zero block (one zero bit);
* fixed rate: every block occupies exactly maxbits bits, so each block decodes alone from its own offset;
* with a budget that holds every bit plane (rate 34) the coder is lossless on the integer coefficients: decoded
  values differ from the input only by the transform's and the cast's rounding (a few units of 2^(emax-30));
* the error falls monotonically with the rate on a smooth field, and partial blocks at the far faces only write
  the valid region.
"""
import numpy as np
import pytest

import zfp_oracle as Z

M = np.array([[4, 4, 4, 4], [5, 1, -1, -5], [-4, 4, 4, -4], [-2, 6, -6, 2]], dtype=np.int64)


def test_lift_equals_published_matrix(rng):
    for _ in range(200):
        v = rng.integers(-(2**14), 2**14, size=4) * 4096
        out = Z.fwd_lift(v)
        assert np.array_equal(out.astype(np.int64) * 16, M @ v.astype(np.int64))
        assert np.array_equal(Z.inv_lift(out), v)


def test_lift_inverse_is_close_on_any_input(rng):
    for _ in range(500):
        v = rng.integers(-(2**29), 2**29, size=4)
        back = Z.inv_lift(Z.fwd_lift(v))
        assert np.max(np.abs(back.astype(np.int64) - v)) <= 4


def test_3d_transform_is_the_matrix_along_x_y_z(rng):
    for _ in range(20):
        blk = (rng.integers(-64, 64, size=(4, 4, 4)) * 4096).astype(np.int64)  # [z, y, x]
        want = np.einsum("ai,bj,ck,kji->cba", M, M, M, blk)  # transform along x, y and z
        got = Z.fwd_xform(blk).reshape(4, 4, 4).astype(np.int64)
        assert np.array_equal(got * 16**3, want)
        assert np.array_equal(Z.inv_xform(got), blk.reshape(64))


def test_negabinary(rng):
    for x in list(rng.integers(-(2**30), 2**30, size=300)) + [0, 1, -1, 2**29, -(2**29)]:
        u = Z.int2uint(int(x))
        assert sum(((u >> i) & 1) * (-2) ** i for i in range(32)) == x
        assert Z.uint2int(u) == x
    assert Z.int2uint(2**29) == 0x60000000 and Z.int2uint(-(2**29)) == 0x20000000


def test_sequency_order():
    p = Z.perm()
    assert sorted(p.tolist()) == list(range(64))
    key = [((q & 3) + ((q >> 2) & 3) + (q >> 4), (q & 3) ** 2 + ((q >> 2) & 3) ** 2 + (q >> 4) ** 2) for q in p]
    assert key == sorted(key)
    assert p[0] == 0 and p[63] == 63


def test_known_answer_constant_blocks():
    assert Z.block_bits(8) == 512
    w = Z.encode_block(np.ones(64), 8)
    # bits: header 1 + 128<<1 (b0, b8); plane 31: 0; plane 30: 1 1 0; plane 29: 1 0; planes 28..0: 0 0
    assert w.tolist() == [0x2D01] + [0] * 7
    w = Z.encode_block(-np.ones(64), 8)
    # plane 31: 0; plane 30: 0; plane 29: 1 1 0; planes 28..0: 0 0
    assert w.tolist() == [0x1901] + [0] * 7
    assert Z.encode_block(np.zeros(64), 8).tolist() == [0] * 8
    for s in (1.0, -1.0):
        f = np.full((4, 4, 4), s, dtype=np.float32)
        out = Z.decompress(Z.compress(f, 8), f.shape, 8)
        assert np.array_equal(out, f)


def test_minimum_block_is_header_only():
    assert Z.block_bits(0.01) == 9
    f = np.full((4, 4, 4), 3.0, dtype=np.float32)
    out = Z.decompress(Z.compress(f, 0.01), f.shape, 0.01)
    assert np.all(out == 0.0)  # no coefficient bits fit


def _smooth(rng, shape):
    z, y, x = np.meshgrid(*[np.linspace(0, 1, n) for n in shape], indexing="ij")
    f = np.sin(6 * x + 1) * np.cos(5 * y) * np.exp(-2 * z) + 0.3 * x * y
    return (f + 1e-3 * rng.standard_normal(shape)).astype(np.float32)


@pytest.mark.parametrize("rate", [4, 8, 12.5, 16])
def test_fixed_rate_random_access(rng, rate):
    f = (rng.standard_normal((9, 10, 11)) * 10.0 ** rng.integers(-3, 3, size=(9, 10, 11))).astype(np.float32)
    words = Z.compress(f, rate)
    bits = Z.block_bits(rate)
    nb = Z.nblocks(11, 10, 9)
    assert words.size == (nb * bits + 63) // 64
    full = Z.decompress(words, f.shape, rate)
    # re-code each block alone and splice it at its offset: the stream is the concatenation
    spliced = np.zeros_like(words)
    b = 0
    pad = lambda i, n: i if n >= 4 else (0 if n == 1 else ((1 if i in (1, 2) else 0) if n == 2 else (0 if i == 3 else i)))
    for bz in range(3):
        for by in range(3):
            for bx in range(3):
                v = np.zeros(64, dtype=np.float32)
                for k in range(4):
                    for j in range(4):
                        for i in range(4):
                            v[i + 4 * j + 16 * k] = f[4 * bz + pad(k, 9 - 4 * bz), 4 * by + pad(j, 10 - 4 * by),
                                                      4 * bx + pad(i, 11 - 4 * bx)]
                one = Z.encode_block(v, rate)
                bitsv = np.unpackbits(one.view(np.uint8), bitorder="little")[:bits]
                allbits = np.unpackbits(spliced.view(np.uint8), bitorder="little")
                allbits[b * bits:(b + 1) * bits] = bitsv
                spliced = np.packbits(allbits, bitorder="little").view(np.uint64)
                b += 1
    assert np.array_equal(spliced, words)
    assert np.all(np.isfinite(full))


def test_lossless_coefficients_at_full_budget(rng):
    f = (rng.standard_normal((8, 8, 8)) * 100).astype(np.float32)
    out = Z.decompress(Z.compress(f, 34), f.shape, 34)
    for bz in range(2):
        for by in range(2):
            for bx in range(2):
                blk = f[4 * bz:4 * bz + 4, 4 * by:4 * by + 4, 4 * bx:4 * bx + 4]
                emax = np.frexp(np.max(np.abs(blk)))[1]
                err = np.abs(out[4 * bz:4 * bz + 4, 4 * by:4 * by + 4, 4 * bx:4 * bx + 4] - blk)
                assert np.max(err) <= 2.0 ** (emax - 30) * 64


def test_error_falls_with_rate(rng):
    f = _smooth(rng, (20, 18, 23))
    errs = []
    for rate in (2, 4, 8, 12, 16, 24):
        out = Z.decompress(Z.compress(f, rate), f.shape, rate)
        errs.append(np.linalg.norm(out - f) / np.linalg.norm(f))
    assert all(a > b for a, b in zip(errs, errs[1:])), errs
    assert errs[3] < 1e-3 and errs[-1] < 1e-6, errs


def test_partial_blocks_write_only_the_valid_region(rng):
    nz, ny, nx, ldx = 7, 6, 13, 20
    f = np.full((nz + 2, ny + 3, ldx), 7.0, dtype=np.float32)
    f[:nz, :ny, :nx] = _smooth(rng, (nz, ny, nx))
    view = f[:nz, :ny, :]
    words = Z.compress(view, 16, nx=nx)
    out = np.full_like(f, -5.0)
    Z.decompress(words, (nz, ny, nx), 16, out=out[:nz, :ny, :])
    assert np.all(out[:nz, :ny, nx:] == -5.0) and np.all(out[nz:] == -5.0) and np.all(out[:, ny:] == -5.0)
    rel = np.linalg.norm(out[:nz, :ny, :nx] - f[:nz, :ny, :nx]) / np.linalg.norm(f[:nz, :ny, :nx])
    assert rel < 1e-4
