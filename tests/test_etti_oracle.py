"""The FP64 eTTI oracle (oracle/etti_oracle.py) pinned by what the mathematics fixes, not by restating its formulas.

* staggered derivatives: with the Taylor order-8 set (1225/1024, -245/3072, 49/5120, -5/7168) the velocity and strain
  operators reproduce the analytic derivatives of degree-7 polynomials at the half points the reference's
  D+ / D- choice implies, times the M22 shift, invd1/invd2 and zprime placement;
* the velocity and strain operators are minus each other's transpose (summation by parts) once their invd2 agree:
  <u, Div sigma> = -sum_k <eps_k(u), sigma_k> over compactly supported fields; a wrong D+/D- pairing or axis fails it;
* interpolation: with the order-8 midpoint weights (1225, -245, 49, -5)/2048 every Mix reproduces p(a + 1/2 - s_a) +
  q(b + 1/2 - s_b) + c (a + 1/2 - s_a)(b + 1/2 - s_b) exactly; the impulse response is diagonal (the same n on both
  axes), which a tensor-product interpolation would not be;
* stress: probing with unit constant strains gives a symmetric 6 x 6 Voigt matrix whose upper triangle, read row by
  row in the order (11, 22, 33, 23, 13, 12), is M1..M21; linear strain fields land at the staggered points each
  interpolation's shifts name (the six precompute launches and the six in-plane mixes);
* mirror, pressure, imaging and the model fill against hand-computed values.
"""
import numpy as np
import pytest

import etti_oracle as E

TAYLOR = (1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168)
MIDPOINT = (1225 / 2048, -245 / 2048, 49 / 2048, -5 / 2048)
N = (13, 11, 9)  # nx, ny, nz
G = E.G


def shape(nx=N[0], ny=N[1], nz=N[2]):
    ldx = ((nx + 8 + 31) // 32) * 32
    return (nz + 8, ny + 8, ldx)


def coords(sh):
    """Array coordinates z, y, x broadcast to the volume shape (grid step 1)."""
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in sh), indexing="ij")
    return z, y, x


def poly(c, s):
    return sum(ck * s ** k for k, ck in enumerate(c))


def dpoly(c, s):
    return sum(k * ck * s ** (k - 1) for k, ck in enumerate(c) if k)


def inner(a, nx=N[0]):
    return E.win(a, nx)


@pytest.fixture
def rng():
    return np.random.default_rng(7)


def _polys(rng, k):
    return [rng.uniform(-1, 1, 8) for _ in range(k)]


def test_velocity_exact_on_polynomials(rng):
    sh = shape()
    z, y, x = coords(sh)
    S = 0.1  # polynomials in s = 0.1 * coordinate keep the values O(1)
    P = _polys(rng, 9)
    f = {}
    f[4] = poly(P[0], S * x)
    f[5] = poly(P[1], S * y)
    f[6] = poly(P[2], S * z)
    f[7] = poly(P[3], S * x) * poly(P[4][:4], S * y)
    f[8] = poly(P[5][:4], S * x) * poly(P[6], S * z)
    f[9] = poly(P[7][:4], S * y) * poly(P[8][:4], S * z)
    M22 = 1.0 + 0.01 * x + 0.02 * y + 0.03 * z
    zprime = rng.uniform(0.5, 1.5, sh[0])
    v = [np.zeros(sh) for _ in range(3)]
    i1, i2 = 1.3, 0.7
    E.velocity(N[0], *v, f[4], f[5], f[6], f[7], f[8], f[9], M22, zprime, TAYLOR, invd=(i1, i2))
    Z, Y, X = (inner(a) for a in (z, y, x))
    zp = zprime[G: G + N[2]][:, None, None]
    m22 = lambda dz, dy, dx: 1.0 + 0.01 * (X + dx) + 0.02 * (Y + dy) + 0.03 * (Z + dz)
    # D- at i - 1/2, D+ at i + 1/2; d/dcoord = S * p'(S * coord)
    r1 = 2 * m22(0, 0, -1) * (S * dpoly(P[0], S * (X - 0.5)) * i1
                              + poly(P[3], S * X) * S * dpoly(P[4][:4], S * (Y + 0.5)) * i2
                              + poly(P[5][:4], S * X) * S * dpoly(P[6], S * (Z + 0.5)) * zp)
    r2 = 2 * m22(0, -1, 0) * (S * dpoly(P[3], S * (X + 0.5)) * poly(P[4][:4], S * Y) * i1
                              + S * dpoly(P[1], S * (Y - 0.5)) * i2
                              + poly(P[7][:4], S * Y) * S * dpoly(P[8][:4], S * (Z + 0.5)) * zp)
    r3 = 2 * m22(-1, 0, 0) * (S * dpoly(P[5][:4], S * (X + 0.5)) * poly(P[6], S * Z) * i1
                              + S * dpoly(P[7][:4], S * (Y + 0.5)) * poly(P[8][:4], S * Z) * i2
                              + S * dpoly(P[2], S * (Z - 0.5)) * zp)
    for got, ref in zip(v, (r1, r2, r3)):
        assert np.allclose(inner(got), ref, rtol=1e-10, atol=1e-12)
        assert np.all(got[:, :, : G] == 0) and np.all(got[: G] == 0)  # ghosts untouched


def test_strain_exact_on_polynomials(rng):
    sh = shape()
    z, y, x = coords(sh)
    S = 0.1
    P = _polys(rng, 6)
    u1 = poly(P[0], S * x) + poly(P[1][:4], S * y) * poly(P[2][:4], S * z)
    u2 = poly(P[3][:4], S * x) * poly(P[4], S * y) + poly(P[5][:4], S * z)
    u3 = poly(P[2], S * x) * poly(P[0][:3], S * y) + poly(P[1], S * z)
    zprime = rng.uniform(0.5, 1.5, sh[0])
    d = [np.zeros(sh) for _ in range(6)]
    i1, i2 = 0.9, 1.7
    E.strain(N[0], u1, u2, u3, *d, zprime, TAYLOR, invd=(i1, i2))
    Z, Y, X = (inner(a) for a in (z, y, x))
    zp = zprime[G: G + N[2]][:, None, None]
    ref = [
        i1 * S * dpoly(P[0], S * (X + 0.5)),
        i2 * S * dpoly(P[1][:4], S * (Y - 0.5)) * poly(P[2][:4], S * Z)
        + i1 * S * dpoly(P[3][:4], S * (X - 0.5)) * poly(P[4], S * Y),
        zp * poly(P[1][:4], S * Y) * S * dpoly(P[2][:4], S * (Z - 0.5))
        + i1 * S * dpoly(P[2], S * (X - 0.5)) * poly(P[0][:3], S * Y),
        i2 * poly(P[3][:4], S * X) * S * dpoly(P[4], S * (Y + 0.5)),
        zp * S * dpoly(P[5][:4], S * (Z - 0.5)) + i2 * poly(P[2], S * X) * S * dpoly(P[0][:3], S * (Y - 0.5)),
        zp * S * dpoly(P[1], S * (Z + 0.5)),
    ]
    for k, (got, r) in enumerate(zip(d, ref)):
        assert np.allclose(inner(got), r, rtol=1e-10, atol=1e-12), k


def test_velocity_is_minus_transpose_of_strain(rng):
    n = 20
    sh = shape(n, n, n)
    sup = np.zeros(sh, dtype=bool)  # support at least 4 points inside the interior on every side
    sup[G + 4: G + n - 4, G + 4: G + n - 4, G + 4: G + n - 4] = True
    rnd = lambda: np.where(sup, rng.standard_normal(sh), 0.0)
    u = [rnd() for _ in range(3)]
    s = {k: rnd() for k in range(4, 10)}
    zprime = np.full(sh[0], 0.7)  # a z-dependent zprime is not part of a transpose pair
    invd = (1.01, 2.01)
    M22 = np.full(sh, 0.5)  # 2 M22 = 1
    v = [np.zeros(sh) for _ in range(3)]
    E.velocity(n, *v, s[4], s[5], s[6], s[7], s[8], s[9], M22, zprime, TAYLOR, invd=invd)
    d = [np.zeros(sh) for _ in range(6)]
    E.strain(n, *u, *d, zprime, TAYLOR, invd=invd)
    dot = lambda a, b: float(np.sum(E.win(a, n) * E.win(b, n)))
    lhs = sum(dot(a, b) for a, b in zip(u, v))
    # strain d4..d9 pairs with stress f4, f7, f8, f5, f9, f6 (d5 = eps_xy pairs with sigma_xy = f7, ...)
    rhs = -(dot(d[0], s[4]) + dot(d[1], s[7]) + dot(d[2], s[8]) + dot(d[3], s[5]) + dot(d[4], s[9])
            + dot(d[5], s[6]))
    assert abs(lhs) > 1.0
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # a mismatch of the y scale (the reference's 2.01 vs 2.02) breaks the identity
    d2 = [np.zeros(sh) for _ in range(6)]
    E.strain(n, *u, *d2, zprime, TAYLOR, invd=(1.01, 2.02))
    rhs2 = -(dot(d2[0], s[4]) + dot(d2[1], s[7]) + dot(d2[2], s[8]) + dot(d2[3], s[5]) + dot(d2[4], s[9])
             + dot(d2[5], s[6]))
    assert abs(lhs - rhs2) > 1e-6 * abs(lhs)


@pytest.mark.parametrize("xs,ys,zs", [(1, 1, 1), (0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 0, 0)])
def test_mixed_exact_on_bilinear_plus_polynomials(xs, ys, zs, rng):
    sh = shape()
    z, y, x = coords(sh)
    S = 0.1
    P = _polys(rng, 3)
    c1, c2 = 0.37, -0.61
    g = poly(P[0], S * x) + poly(P[1], S * y) + poly(P[2], S * z) + c1 * x * z + c2 * y * z
    xz, yz = np.zeros(sh), np.zeros(sh)
    E.mixed_xz_yz(N[0], g, xz, yz, xs, ys, zs, MIDPOINT)
    Z, Y, X = (inner(a) for a in (z, y, x))
    xp, yp, zq = X + 0.5 - xs, Y + 0.5 - ys, Z + 0.5 - zs
    ref_xz = poly(P[0], S * xp) + poly(P[1], S * Y) + poly(P[2], S * zq) + c1 * xp * zq + c2 * Y * zq
    ref_yz = poly(P[0], S * X) + poly(P[1], S * yp) + poly(P[2], S * zq) + c1 * X * zq + c2 * yp * zq
    assert np.allclose(inner(xz), ref_xz, rtol=1e-11, atol=1e-11)
    assert np.allclose(inner(yz), ref_yz, rtol=1e-11, atol=1e-11)


def test_mixed_impulse_response_is_diagonal():
    sh = shape()
    si = (0.1, 0.2, 0.3, 0.4)
    g = np.zeros(sh)
    x0, z0 = G + 7, G + 4
    g[z0, G + 5, x0] = 1.0
    xz, yz = np.zeros(sh), np.zeros(sh)
    xs, zs = 1, 0
    E.mixed_xz_yz(N[0], g, xz, yz, xs, 0, zs, si)
    at = lambda a, b: xz[z0 - b + zs, G + 5, x0 - a + xs]  # output whose (p, q) = (a, b) reads the impulse
    assert at(4, 4) == 0.5 * si[3] and at(-3, 4) == 0.5 * si[3] and at(4, -3) == 0.5 * si[3]
    assert at(1, 0) == 0.5 * si[0] and at(2, -1) == 0.5 * si[1]
    assert at(4, 1) == 0.0 and at(1, 4) == 0.0 and at(2, 3) == 0.0  # different n on the two axes
    assert np.count_nonzero(xz) == 16


def _stress_inputs(sh, dvals, mvals):
    f = {"f%d" % k: np.zeros(sh) for k in range(4, 22)}
    d = {"d%d" % k: (v if isinstance(v, np.ndarray) else np.full(sh, float(v))) for k, v in dvals.items()}
    M = [None] + [np.full(sh, float(mvals[i])) for i in range(1, 22)]
    return f, d, M


def test_stress_voigt_matrix_symmetric_and_ordered():
    sh = shape()
    rows = ["f4", "f5", "f6", "f9", "f8", "f7"]  # sigma 11, 22, 33, 23, 13, 12
    cols = [4, 7, 9, 8, 6, 5]  # eps 11, 22, 33, 23, 13, 12
    C = np.zeros((6, 6))
    for j, dk in enumerate(cols):
        dv = {k: (1.0 if k == dk else 0.0) for k in range(4, 10)}
        f, d, M = _stress_inputs(sh, dv, list(range(22)))  # M_k = k
        E.stress_loop(N[0], f, M, d, MIDPOINT)
        for i, r in enumerate(rows):
            vals = inner(f[r])
            assert np.allclose(vals, vals.flat[0], rtol=0, atol=1e-13)
            C[i, j] = vals.flat[0]
    assert np.allclose(C, C.T, atol=1e-12)
    assert np.allclose(C[np.triu_indices(6)], np.arange(1, 22), atol=1e-12)


def test_stress_interpolation_points():
    sh = shape()
    z, y, x = coords(sh)
    grads = {4: (1.0, 0.3, 0.2), 5: (0.5, -0.7, 0.1), 6: (-0.2, 0.4, 0.9), 7: (0.8, 0.1, -0.3),
             8: (0.25, 0.6, -0.45), 9: (-0.55, 0.35, 0.15)}
    lin = lambda k, dz=0.0, dy=0.0, dx=0.0: grads[k][0] * (x + dx) + grads[k][1] * (y + dy) + grads[k][2] * (z + dz)
    d = {"d%d" % k: lin(k) for k in range(4, 10)}
    # use one model entry at a time to read off each term
    terms = {}
    for mi in range(1, 22):
        f2, d2, M2 = _stress_inputs(sh, {k: d["d%d" % k] for k in range(4, 10)}, [0] + [0.0] * 21)
        M2[mi] = np.ones(sh)
        E.stress_loop(N[0], f2, M2, d2, MIDPOINT)
        terms[mi] = {r: inner(f2[r]).copy() for r in ("f4", "f5", "f6", "f7", "f8", "f9")}
    L = lambda k, dz=0.0, dy=0.0, dx=0.0: inner(lin(k, dz, dy, dx))
    # (model, stress) -> strain sampled where the interpolation puts it: a mix over axes (a, b) with shifts
    # (sa, sb) samples at a + 1/2 - sa, b + 1/2 - sb
    expect = {
        (1, "f4"): L(4), (2, "f4"): L(7), (3, "f4"): L(9),
        (6, "f4"): L(5, 0, 0.5, 0.5),     # xy mix of d5, shifts (0, 0)
        (5, "f4"): L(6, 0.5, 0, 0.5),     # f14: xz of d6, XS = ZS = 0
        (4, "f4"): L(8, 0.5, 0.5, 0),     # f19: yz of d8, YS = ZS = 0
        (21, "f7"): L(5),
        (6, "f7"): L(4, 0, -0.5, -0.5),   # xy mix of d4, shifts (1, 1)
        (11, "f7"): L(7, 0, -0.5, -0.5), (15, "f7"): L(9, 0, -0.5, -0.5),
        (20, "f7"): L(6, 0.5, -0.5, 0),   # f15: yz of d6, YS = 1, ZS = 0
        (18, "f7"): L(8, 0.5, 0, -0.5),   # f18: xz of d8, XS = 1, ZS = 0
        (19, "f8"): L(6),
        (5, "f8"): L(4, -0.5, 0, -0.5),   # f10: xz of d4, XS = ZS = 1
        (10, "f8"): L(7, -0.5, 0, -0.5), (14, "f8"): L(9, -0.5, 0, -0.5),
        (20, "f8"): L(5, -0.5, 0.5, 0),   # f13: yz of d5, YS = 0, ZS = 1
        (17, "f8"): L(8, 0, 0.5, -0.5),   # xy mix of d8, shifts (1, 0)
        (16, "f9"): L(8),
        (4, "f9"): L(4, -0.5, -0.5, 0),   # f11: yz of d4, YS = ZS = 1
        (9, "f9"): L(7, -0.5, -0.5, 0), (13, "f9"): L(9, -0.5, -0.5, 0),
        (18, "f9"): L(5, -0.5, 0, 0.5),   # f12: xz of d5, XS = 0, ZS = 1
        (17, "f9"): L(6, 0, -0.5, 0.5),   # xy mix of d6, shifts (0, 1)
    }
    for (mi, r), ref in expect.items():
        assert np.allclose(terms[mi][r], ref, rtol=1e-12, atol=1e-12), (mi, r)


def test_mirror_pressure_imaging_and_model_fill():
    sh = shape()
    a = np.arange(np.prod(sh), dtype=np.float64).reshape(sh) + 1.0
    b = a.copy()
    E.mirror([b])
    assert np.all(b[G] == 0)
    assert np.array_equal(b[G - 1], -a[G + 1]) and np.array_equal(b[0], -a[2 * G])
    assert np.array_equal(b[G + 1:], a[G + 1:])
    s = np.full(sh, 2.0)
    assert np.all(E.compute_p(s, s, s, 0 * s, 0 * s, 0 * s) == -2.0 / 3.0)  # -(2 + 3 + 4) * 2 / 27
    assert E.compute_p(*(np.array([1.0 * k]) for k in range(1, 7)))[0] == -(2 + 6 + 12 + 20 + 30 + 42) / 27
    nx, ny, nz = N
    g1 = np.ones((nz, ny, nx))
    a1 = np.zeros(sh)
    a3 = np.zeros(sh)
    a1[G + 1, G + 2, G + 3] = 2.0
    a3[G + 1, G + 2, G + 3] = 5.0
    s1 = np.full((nz, ny, nx), 3.0)
    s3 = np.full((nz, ny, nx), -1.0)
    E.update_gv(nx, g1, s1, s3, a1, a3)
    assert g1[1, 2, 3] == 1.0 + 6.0 - 5.0 and np.count_nonzero(g1 != 1.0) == 1
    vols, zprime = E.init_cijs(sh, nx)
    assert vols["M1"][0, 0, G] == np.cos(4.0) and vols["M22"][2, 1, G + 1] == np.sin(2 + 5 + 5)
    assert vols["d9"][2, 1, G] == np.sin(2 + 6 + 4) and vols["d6"][3, 0, G + 2] == np.cos(18 + 0 + 6)
    assert vols["M11"][1, 2, G + nx - 1] == np.cos(1 + 2 + 3 * (G + nx - 1))
    assert np.all(vols["M7"][:, :, :G] == 0) and np.all(vols["M7"][:, :, G + nx:] == 0)
    assert zprime[5] == np.cos(30.0)


def test_driver_exchange_fills_ghosts_from_neighbours():
    # 3 x 2 x 3 zones of uneven sizes cut from one global field: after the X -> Y -> Z exchange every ghost that has a
    # neighbour holds the global field there (edges and corners included); ghosts on the global boundary stay zero
    sizes = ([5, 9, 6], [7, 4], [6, 5, 8])
    offs = [np.concatenate([[0], np.cumsum(s)[:-1]]) for s in sizes]
    tot = [sum(s) for s in sizes]
    glob = lambda z, y, x: 1.0 + x + 100.0 * y + 10000.0 * z
    subs = []
    for iz, (oz, nz) in enumerate(zip(offs[2], sizes[2])):
        for iy, (oy, ny) in enumerate(zip(offs[1], sizes[1])):
            for ix, (ox, nx) in enumerate(zip(offs[0], sizes[0])):
                a = np.zeros(shape(nx, ny, nz))
                z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
                a[G: G + nz, G: G + ny, G: G + nx] = glob(z + oz, y + oy, x + ox)
                subs.append({"offset": (ox, oy, oz), "n": (nx, ny, nz), "vols": {"f4": a}, "zprime": None})
    sim = E.Sim3(subs, TAYLOR, MIDPOINT)
    sim.exchange(["f4"])
    for s in subs:
        (ox, oy, oz), (nx, ny, nz) = s["offset"], s["n"]
        a = s["vols"]["f4"]
        z, y, x = np.meshgrid(np.arange(nz + 8) - G + oz, np.arange(ny + 8) - G + oy, np.arange(nx + 8) - G + ox,
                              indexing="ij")
        ref = glob(z, y, x)
        inside = (x >= 0) & (x < tot[0]) & (y >= 0) & (y < tot[1]) & (z >= 0) & (z < tot[2])
        got = a[:, :, : nx + 8]
        assert np.array_equal(got[inside], ref[inside])
        assert not np.any(got[~inside])
