"""Pins the FP64 oracle against what mathematics fixes, independently of its own formulas.

* coefficients: the staggered set equals the Taylor set 1225/1024, -245/3072, 49/5120, -5/7168 to FP32 rounding; the
  central set satisfies consistency (e0 + 2 sum e_n ~ 0) and scale (sum n^2 e_n ~ 1);
* every derivative operator reproduces its plane-wave symbol (a wrong index, sign or coefficient changes the phase
  or the amplitude of the symbol);
* TTI invariants: isotropy when eps = delta = vs = 0 for any tilt; VTI (tilt 0) has V = F_zz and no azimuth
  dependence; tilt 90 deg about y equals the tilt-0 result with x and z swapped;
* the adjoint operators are the exact transposes of the forward ones (dot-product tests in FP64), for one operator
  application and for a multi-step propagation;
* the density gradient equals minus the derivative of <mu, L(rho) sigma> with respect to rho (central differences);
* pointwise kernels (mirrors, model build) satisfy their defining symmetries;
* the imaging condition lands each impulse at its own no-halo index and is bilinear; its finalisation is -2 dt_xcorr g
  at vp = 1 and scales as 1/vp (a missing square root or factor fails); pressure is minus the mean normal stress of
  diag(sigma_H, sigma_H, sigma_V).
"""
import numpy as np
import pytest

import oracle as O

DX, DY, DZ = 10.0, 12.5, 8.0


def plane(G, k, phase=0.3):
    """cos(kx*x + ky*y + kz*z + phase) on every array point (ghosts included), x = ix*dx etc."""
    z, y, x = np.meshgrid(np.arange(G.ldz) * DZ, np.arange(G.ldy) * DY, np.arange(G.ldx) * DX, indexing="ij")
    return np.cos(k[0] * x + k[1] * y + k[2] * z + phase), (x, y, z)


def stag_symbol(theta):
    c, _ = O.coefs()
    return 2 * sum(c[n] * np.sin((n - 0.5) * theta) for n in range(1, 5))


def cen_symbol(theta):
    _, e = O.coefs()
    return e[0] + 2 * sum(e[n] * np.cos(n * theta) for n in range(1, 5))


def mixed_symbol(ta, tb):
    _, e = O.coefs()
    return -sum(e[n] * np.sin(n * ta) * np.sin(n * tb) for n in range(1, 5))


def test_coefficients_closed_form():
    c, e = O.coefs()
    taylor = [1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168]
    for n in range(1, 5):
        assert abs(c[n] - taylor[n - 1]) <= 1e-7 * abs(taylor[n - 1])
    assert abs(e[0] + 2 * e[1:].sum()) < 1e-6
    assert abs(sum(n * n * e[n] for n in range(1, 5)) - 1.0) < 1e-5
    # staggered scheme: consistency 2 sum (n-1/2) c_n = 1 and exactness on cubics (Taylor coefficients)
    assert abs(2 * sum((n - 0.5) * c[n] for n in range(1, 5)) - 1.0) < 1e-6
    assert abs(2 * sum((n - 0.5) ** 3 * c[n] for n in range(1, 5))) < 1e-6


def test_staggered_symbol_accuracy():
    # the 8th-order staggered symbol matches the exact one at low wavenumber
    for th in (0.05, 0.2, 0.5):
        assert abs(stag_symbol(th) / th - 1.0) < 5e-6 * (th / 0.05) ** 8 + 1e-6


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fwd_drv_plane_wave(axis):
    G = O.Geom(20, 18, 17)
    k = [0.0, 0.0, 0.0]
    h = (DX, DY, DZ)[axis]
    k[axis] = 1.3 / h
    f, (x, y, z) = plane(G, k)
    b = np.full(G.shape, 0.7)
    d = [G.zeros() for _ in range(6)]
    c = O.fd_consts(DX, DY, DZ)
    O.fwd_drv(G, f, 2 * f, b, d, 1 / DX, 1 / DY, 1 / DZ)
    # D- cos(phi)(i) = -(S/h) sin(phi(i - 1/2))
    shift = [0.0, 0.0, 0.0]
    shift[axis] = -0.5 * h
    expect = -0.7 * stag_symbol(k[axis] * h) / h * np.sin(k[0] * (x + shift[0]) + k[1] * (y + shift[1]) + k[2] * (z + shift[2]) + 0.3)
    got = d[axis]
    got2 = d[3 + axis]
    err = np.abs(G.interior(got) - G.interior(expect)).max() / np.abs(G.interior(expect)).max()
    err2 = np.abs(G.interior(got2) - 2 * G.interior(expect)).max() / np.abs(G.interior(expect)).max()
    assert err < 1e-12 and err2 < 1e-12
    # ghosts are never written
    assert np.all(got[:4] == 0) and np.all(got[:, :4] == 0) and np.all(got[:, :, :4] == 0)
    del c


def deep(G, a, r=4):
    """Interior shrunk by r: points whose D+ stencil does not reach the (zero) derivative ghosts."""
    g = G.g + r
    return a[g : G.ldz - g, g : G.ldy - g, g : g + G.nx - 2 * r]


def _models_const(G, vp, eps, delta, vs, theta, phi, rho):
    full = lambda v: np.full(G.shape, float(v))
    m = O.init_cijs(full(vp), full(eps), full(delta), full(vs), full(theta), full(phi), full(rho))
    return m  # m1..m7, b


def _fwd_rhs(G, s1, s2, models, dx=DX, dy=DY, dz=DZ):
    """rhs = L(s) via the oracle's derivative + update operators (u starts at zero, dt = 1)."""
    d = [G.zeros() for _ in range(6)]
    c = O.fd_consts(dx, dy, dz)
    O.fwd_drv(G, s1, s2, models[7], d, c["invdx"], c["invdy"], c["invdz"])
    u, w = G.zeros(), G.zeros()
    O.fwd_update(G, u, w, d, s1, s2, models, 1.0, c["invdx"], c["invdy"], c["invdz"], c["invdxy"], c["invdyz"],
                 c["invdzx"])
    return u, w


def _symbols(k, dx=DX, dy=DY, dz=DZ):
    fxx = -(stag_symbol(k[0] * dx) / dx) ** 2
    fyy = -(stag_symbol(k[1] * dy) / dy) ** 2
    fzz = -(stag_symbol(k[2] * dz) / dz) ** 2
    fxy = mixed_symbol(k[0] * dx, k[1] * dy) / (dx * dy)
    fyz = mixed_symbol(k[1] * dy, k[2] * dz) / (dy * dz)
    fzx = mixed_symbol(k[2] * dz, k[0] * dx) / (dz * dx)
    return fxx, fyy, fzz, fxy, fyz, fzx


@pytest.mark.parametrize("theta,phi", [(0.0, 0.0), (0.4, 1.1), (1.2, 4.0)])
def test_isotropic_invariant(theta, phi):
    """eps = delta = vs = 0 and sigmaH = sigmaV: rhs3 = rhs4 = vp^2 * discrete Laplacian for every tilt."""
    G = O.Geom(22, 20, 21)
    k = (0.9 / DX, 0.6 / DY, 1.1 / DZ)
    f, _ = plane(G, k)
    m = _models_const(G, 2000.0, 0.0, 0.0, 0.0, theta, phi, 1800.0)
    u, w = _fwd_rhs(G, f, f, m)
    fxx, fyy, fzz, *_ = _symbols(k)
    expect = 2000.0**2 * (fxx + fyy + fzz) * f
    for got in (u, w):
        err = np.abs(deep(G, got - expect)).max() / np.abs(deep(G, expect)).max()
        assert err < 1e-6  # FP32-rounded 1/dx constants (relative 6e-8 each)


def _tti_symbol_rhs(k, m_scalar, r):
    m1, m2, m3, m4 = m_scalar
    fxx, fyy, fzz, fxy, fyz, fzx = _symbols(k)
    rx, ry, rz = r
    V = rx * rx * fxx + ry * ry * fyy + rz * rz * fzz + 2 * rx * ry * fxy + 2 * ry * rz * fyz + 2 * rz * rx * fzx
    H = fxx + fyy + fzz - V
    return V, H


@pytest.mark.parametrize("theta,phi", [(0.0, 0.7), (0.5, 2.0), (1.3, 5.5)])
def test_tti_plane_wave(theta, phi):
    """Full TTI operator on plane waves with sigmaH and sigmaV of different amplitude vs its symbol."""
    G = O.Geom(21, 23, 20)
    k = (0.7 / DX, -0.5 / DY, 0.9 / DZ)
    f, _ = plane(G, k)
    vp, eps, delta, vs, rho = 2500.0, 0.25, 0.1, 600.0, 2100.0
    m = _models_const(G, vp, eps, delta, vs, theta, phi, rho)
    a1, a2 = 1.0, -0.6
    u, w = _fwd_rhs(G, a1 * f, a2 * f, m)
    m1, m2, m3, m4 = (float(x.flat[0]) for x in m[:4])
    r = (np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta))
    V, H = _tti_symbol_rhs(k, (m1, m2, m3, m4), r)
    e3 = (m1 * H * a1 + m4 * V * a1 + m2 * V * a2) * f
    e4 = (m2 * H * a1 + m3 * V * a2 + m4 * H * a2) * f
    for got, ex in ((u, e3), (w, e4)):
        err = np.abs(deep(G, got - ex)).max() / np.abs(deep(G, ex)).max()
        assert err < 1e-6  # FP32-rounded 1/dx constants (relative 6e-8 each)


def test_vti_azimuth_independence_and_v_is_fzz():
    G = O.Geom(20, 20, 20)
    rng = np.random.default_rng(1)
    s1 = rng.standard_normal(G.shape)
    s2 = rng.standard_normal(G.shape)
    outs = []
    for phi in (0.0, 1.0, 3.0):
        m = _models_const(G, 3000.0, 0.2, 0.1, 0.0, 0.0, phi, 2000.0)
        outs.append(_fwd_rhs(G, s1, s2, m))
    for u, w in outs[1:]:
        assert np.abs(G.interior(u - outs[0][0])).max() <= 1e-12 * np.abs(G.interior(outs[0][0])).max()
        assert np.abs(G.interior(w - outs[0][1])).max() <= 1e-12 * np.abs(G.interior(outs[0][1])).max()


def test_tilt90_equals_permuted_vti():
    """theta = 90 deg, phi = 0 (symmetry axis along x) equals the VTI operator with x and z exchanged (cube grid)."""
    n = 20
    G = O.Geom(n, n, n, align=1)
    rng = np.random.default_rng(2)
    s1 = rng.standard_normal(G.shape)
    s2 = rng.standard_normal(G.shape)
    h = 10.0
    m_tti = _models_const(G, 2800.0, 0.2, 0.08, 300.0, np.pi / 2, 0.0, 1900.0)
    m_vti = _models_const(G, 2800.0, 0.2, 0.08, 300.0, 0.0, 0.0, 1900.0)
    # cos(pi/2) is not exactly 0 in floating point; the symmetry is checked to that level
    u1, w1 = _fwd_rhs(G, s1, s2, m_tti, h, h, h)
    sw = lambda a: np.ascontiguousarray(np.swapaxes(a, 0, 2))
    u2, w2 = _fwd_rhs(G, sw(s1), sw(s2), m_vti, h, h, h)
    for a, b in ((u1, sw(u2)), (w1, sw(w2))):
        assert np.abs(G.interior(a - b)).max() < 1e-9 * np.abs(G.interior(a)).max()


def _random_models(G, rng):
    vp = rng.uniform(1500, 4500, G.shape)
    eps = rng.uniform(0, 0.3, G.shape)
    delta = rng.uniform(0, 0.15, G.shape)
    vs = rng.uniform(0, 0.4, G.shape) * vp
    theta = rng.uniform(0, np.pi / 3, G.shape)
    phi = rng.uniform(0, 2 * np.pi, G.shape)
    rho = rng.uniform(1000, 2600, G.shape)
    return O.init_cijs(vp, eps, delta, vs, theta, phi, rho)


def _zero_ghost(G, a):
    out = G.zeros()
    G.interior(out)[...] = G.interior(a)
    return out


def _adj_rhs(G, a1, a2, models, dx=DX, dy=DY, dz=DZ):
    d = [G.zeros() for _ in range(6)]
    r3, r4 = G.zeros(), G.zeros()
    c = O.fd_consts(dx, dy, dz)
    O.adj_drv(G, d, r3, r4, a1, a2, models, c["invdx"], c["invdy"], c["invdz"], c["invdxy"], c["invdyz"], c["invdzx"])
    u, w = G.zeros(), G.zeros()
    O.adj_update(G, u, w, d, r3, r4, 1.0, c["invdx"], c["invdy"], c["invdz"])
    return u, w


def test_type2_adjoint_is_transpose():
    G = O.Geom(17, 15, 16)
    rng = np.random.default_rng(3)
    m = _random_models(G, rng)
    s1, s2, a1, a2 = (_zero_ghost(G, rng.standard_normal(G.shape)) for _ in range(4))
    Lu, Lw = _fwd_rhs(G, s1, s2, m)
    Tu, Tw = _adj_rhs(G, a1, a2, m)
    lhs = np.sum(G.interior(Lu) * G.interior(a1)) + np.sum(G.interior(Lw) * G.interior(a2))
    rhs = np.sum(G.interior(s1) * G.interior(Tu)) + np.sum(G.interior(s2) * G.interior(Tw))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))


def test_type1_adjoint_is_transpose():
    G = O.Geom(16, 17, 15)
    rng = np.random.default_rng(4)
    m = _random_models(G, rng)[:7]
    c = O.fd_consts(DX, DY, DZ)
    s1, s2, a1, a2 = (_zero_ghost(G, rng.standard_normal(G.shape)) for _ in range(4))
    Lu, Lw = G.zeros(), G.zeros()
    O.t1_fwd(G, Lu, Lw, s1, s2, m, 1.0, c)
    Tu, Tw = G.zeros(), G.zeros()
    O.t1_adj(G, Tu, Tw, a1, a2, m, 1.0, c)
    lhs = np.sum(G.interior(Lu) * G.interior(a1)) + np.sum(G.interior(Lw) * G.interior(a2))
    rhs = np.sum(G.interior(s1) * G.interior(Tu)) + np.sum(G.interior(s2) * G.interior(Tw))
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs))


def test_type1_central_symbol():
    G = O.Geom(20, 20, 20)
    k = (0.8 / DX, 0.4 / DY, 1.0 / DZ)
    f, _ = plane(G, k)
    m = _models_const(G, 2000.0, 0.0, 0.0, 0.0, 0.3, 0.2, 1000.0)[:7]
    c = O.fd_consts(DX, DY, DZ)
    u, w = G.zeros(), G.zeros()
    O.t1_fwd(G, u, w, f, f, m, 1.0, c)
    lap = cen_symbol(k[0] * DX) / DX**2 + cen_symbol(k[1] * DY) / DY**2 + cen_symbol(k[2] * DZ) / DZ**2
    ex = 2000.0**2 * lap * f
    assert np.abs(G.interior(u - ex)).max() < 1e-6 * np.abs(G.interior(ex)).max()


def test_propagator_dot_product():
    """N forward steps vs N adjoint steps (A = B = 1, zero ghosts): <P x, y> = <x, P^T y> in FP64."""
    G = O.Geom(16, 14, 15)
    rng = np.random.default_rng(5)
    m = _random_models(G, rng)
    sp = O.unit_sponge(G)
    dt = 4e-4
    x = [_zero_ghost(G, rng.standard_normal(G.shape)) for _ in range(4)]  # s1 s2 u w
    y = [_zero_ghost(G, rng.standard_normal(G.shape)) for _ in range(4)]  # lambda(s1) lambda(s2) mu(u) mu(w)
    f1, f2, u, w = (a.copy() for a in x)
    d = [G.zeros() for _ in range(6)]
    for _ in range(6):
        O.step_fwd(G, f1, f2, u, w, m, d, sp, dt, DX, DY, DZ)
    # adjoint state in the reference labelling: sigma_adj = dual of u, u_adj = dual of sigma
    as1, as2, au, aw = y[2].copy(), y[3].copy(), y[0].copy(), y[1].copy()
    d8 = [G.zeros() for _ in range(8)]
    for _ in range(6):
        O.step_adj(G, as1, as2, au, aw, m, d8, sp, dt, DX, DY, DZ)
    I = G.interior
    lhs = sum(np.sum(I(a) * I(b)) for a, b in ((f1, y[0]), (f2, y[1]), (u, y[2]), (w, y[3])))
    rhs = np.sum(I(x[0]) * I(au)) + np.sum(I(x[1]) * I(aw)) + np.sum(I(x[2]) * I(as1)) + np.sum(I(x[3]) * I(as2))
    assert abs(lhs - rhs) <= 1e-11 * max(abs(lhs), abs(rhs))


def test_rho_gradient_matches_finite_difference():
    """rho_gradient(x) = -d/drho(x) <mu, L(rho) sigma>; checked along a random direction with central differences."""
    G = O.Geom(14, 13, 15)
    rng = np.random.default_rng(6)
    vp = rng.uniform(1800, 3500, G.shape)
    eps = rng.uniform(0, 0.3, G.shape)
    delta = rng.uniform(0, 0.15, G.shape)
    vs = rng.uniform(0, 0.3, G.shape) * vp
    th = rng.uniform(0, 1.0, G.shape)
    ph = rng.uniform(0, 6.28, G.shape)
    rho = rng.uniform(1200, 2400, G.shape)
    s1, s2, mu1, mu2 = (_zero_ghost(G, rng.standard_normal(G.shape)) for _ in range(4))
    direction = _zero_ghost(G, rng.standard_normal(G.shape))

    def J(r):
        m = O.init_cijs(vp, eps, delta, vs, th, ph, r)
        u, w = _fwd_rhs(G, s1, s2, m)
        return np.sum(G.interior(u) * G.interior(mu1)) + np.sum(G.interior(w) * G.interior(mu2))

    h = 1e-3
    fd = (J(rho + h * direction) - J(rho - h * direction)) / (2 * h)
    m = O.init_cijs(vp, eps, delta, vs, th, ph, rho)
    c = O.fd_consts(DX, DY, DZ)
    fwd_d = [G.zeros() for _ in range(6)]
    O.fwd_drv(G, s1, s2, m[7], fwd_d, c["invdx"], c["invdy"], c["invdz"])
    grad = np.zeros((G.nz, G.ny, G.nx))
    O.rho_gradient(G, grad, fwd_d, mu1, mu2, m, c["invdx"], c["invdy"], c["invdz"], G.nx, G.ny)
    an = -np.sum(grad * G.interior(direction))
    assert abs(an - fd) <= 1e-7 * abs(fd)


def test_sigma_mirror_is_odd():
    G = O.Geom(12, 11, 13)
    rng = np.random.default_rng(7)
    f1, f2 = rng.standard_normal(G.shape), rng.standard_normal(G.shape)
    O.loop4(G, f1, f2)
    g = G.g
    assert np.all(f1[g] == 0) and np.all(f2[g] == 0)
    for k in range(1, g + 1):
        assert np.array_equal(f1[g - k], -f1[g + k]) and np.array_equal(f2[g - k], -f2[g + k])


def test_derivative_mirror_symmetries():
    G = O.Geom(12, 11, 13)
    rng = np.random.default_rng(8)
    d = [rng.standard_normal(G.shape) for _ in range(6)]
    before = [a.copy() for a in d]
    O.loop4_derivatives(G, d)
    g = G.g
    for a in (0, 1, 3, 4):  # x and y derivatives: odd about izbeg, zero on it
        assert np.all(d[a][g] == 0)
        for k in range(1, g + 1):
            assert np.array_equal(d[a][g - k], -d[a][g + k])
    for a in (2, 5):  # z derivatives: even about izbeg + 1/2
        for k in range(0, g + 1):
            assert np.array_equal(d[a][g - k], d[a][g + 1 + k])
        assert np.array_equal(d[a][g + 1 :], before[a][g + 1 :])


def test_init_cijs_identities():
    rng = np.random.default_rng(9)
    n = 1000
    vp = rng.uniform(1500, 4500, n)
    eps, delta = rng.uniform(0, 0.3, n), rng.uniform(0, 0.15, n)
    vs = rng.uniform(0, 0.5, n) * vp
    th, ph, rho = rng.uniform(0, 1.5, n), rng.uniform(0, 6.3, n), rng.uniform(1000, 3000, n)
    m1, m2, m3, m4, m5, m6, m7, b = O.init_cijs(vp, eps, delta, vs, th, ph, rho)
    assert np.allclose(m1 / m3, 1 + 2 * eps, rtol=1e-14)
    assert np.allclose(m2**2, (m3 - m4) * (m3 * (1 + 2 * delta) - m4), rtol=1e-13)
    assert np.allclose(m5**2 + m6**2 + m7**2, 1.0, rtol=1e-14)
    assert np.allclose(b * rho, 1.0, rtol=1e-15)
    assert np.allclose(m4, vs**2, rtol=1e-15)


def test_loop2_sponge_min_rule():
    G = O.Geom(9, 8, 7)
    rng = np.random.default_rng(10)
    f1, f2, u, w = (rng.standard_normal(G.shape) for _ in range(4))
    abx = rng.uniform(0.5, 1, (G.ldx, 2)).astype(np.float32)
    aby = rng.uniform(0.5, 1, (G.ldy, 2)).astype(np.float32)
    abz = rng.uniform(0.5, 1, (G.ldz, 2)).astype(np.float32)
    f1_0 = f1.copy()
    O.loop2(G, f1, f2, u, w, abx, aby, abz, 0.01)
    iz, iy, ix = 5, 6, 7
    a = min(abx[ix, 0], aby[iy, 0], abz[iz, 0])
    b = min(abx[ix, 1], aby[iy, 1], abz[iz, 1])
    assert abs(f1[iz, iy, ix] - (a * f1_0[iz, iy, ix] + b * u[iz, iy, ix] * 0.01)) < 1e-14
    assert f1[0, 0, 0] == f1_0[0, 0, 0]  # ghosts untouched


def test_vp_gradient_layout_and_bilinearity():
    """Imaging condition (kernels/kernel_type2_small.cu:42-59): the adjoint fields are read at the halo'd index of an
    interior point and the snapshots and the image at its no-halo index. A unit impulse in one adjoint field at one
    interior point, against constant snapshots, must land at exactly that point's no-halo index, paired with its own
    snapshot field (u with snap3, w with snap4); the update accumulates and is bilinear."""
    G = O.Geom(9, 7, 6)
    pt = (5, 3, 2)  # interior (ix, iy, iz) offsets
    i = (G.g + pt[2], G.g + pt[1], G.g + pt[0])
    for which, sv in ((0, 2.0), (1, 5.0)):
        adj = [G.zeros(), G.zeros()]
        adj[which][i] = 1.0
        snap3, snap4 = np.full((G.nz, G.ny, G.nx), 2.0), np.full((G.nz, G.ny, G.nx), 5.0)
        grad = np.full((G.nz, G.ny, G.nx), 0.25)
        O.vp_gradient(G, snap3, snap4, grad, adj[0], adj[1])
        expect = np.full_like(grad, 0.25)
        expect[pt[2], pt[1], pt[0]] += sv
        assert np.array_equal(grad, expect)
    rng = np.random.default_rng(12)
    a3, a4 = rng.standard_normal(G.shape), rng.standard_normal(G.shape)
    s3, s4 = rng.standard_normal((G.nz, G.ny, G.nx)), rng.standard_normal((G.nz, G.ny, G.nx))

    def img(a3, a4, s3, s4):
        g = np.zeros((G.nz, G.ny, G.nx))
        O.vp_gradient(G, s3, s4, g, a3, a4)
        return g

    base = img(a3, a4, s3, s4)
    assert np.allclose(img(3 * a3, 3 * a4, s3, s4), 3 * base, rtol=1e-14, atol=1e-14)
    assert np.allclose(img(a3, a4, -2 * s3, -2 * s4), -2 * base, rtol=1e-14, atol=1e-14)
    assert np.allclose(img(a3, 0 * a4, s3, s4) + img(0 * a3, a4, s3, s4), base, rtol=1e-14, atol=1e-14)


def test_finalize_vp_gradient_properties():
    """finalize_Vp_gradient (kernels/kernel_type2_small.cu:61-76) scales the image by -2 dt_xcorr / vp with
    vp = sqrt(model3). Pinned without its formula: at vp = 1 the result is exactly -2 dt_xcorr g (sign and factor 2);
    scaling model3 by s^2 scales the result by 1/s (a missing or doubled square root fails); the map is linear in g;
    a model3 of 4 at a single interior point halves exactly that point's output (model read at the halo'd index of the
    point whose no-halo image value it scales)."""
    G = O.Geom(8, 6, 5)
    rng = np.random.default_rng(13)
    g = rng.uniform(0.5, 2.0, (G.nz, G.ny, G.nx))
    dtx = 0.004

    def fin(model3, g):
        out = g.copy()
        O.finalize_vp_gradient(G, model3, out, dtx)
        return out

    ones = np.ones(G.shape)
    r1 = fin(ones, g)
    assert np.array_equal(r1, -2.0 * dtx * g)
    assert (r1 < 0).all()
    for s in (2.0, 10.0, 1500.0):
        assert np.allclose(fin(s * s * ones, g), r1 / s, rtol=1e-14)
    g2 = rng.standard_normal(g.shape)
    assert np.allclose(fin(ones * 9.0, 2 * g - 3 * g2), 2 * fin(ones * 9.0, g) - 3 * fin(ones * 9.0, g2), rtol=1e-13,
                       atol=1e-15)
    m = ones.copy()
    pt = (6, 1, 3)
    m[G.g + pt[2], G.g + pt[1], G.g + pt[0]] = 4.0
    r = fin(m, g)
    expect = r1.copy()
    expect[pt[2], pt[1], pt[0]] *= 0.5
    assert np.array_equal(r, expect)


def test_pressure_is_minus_mean_normal_stress():
    """Pressure of simulator2/simulator2SetUp.cu:31-42 is minus the mean normal stress of the TTI stress state
    diag(sigma_H, sigma_H, sigma_V): a hydrostatic state sigma_H = sigma_V = s gives -s, the two horizontal components
    weigh 2/3 and the vertical 1/3 (a swapped weight fails the asymmetric cases), and the map is linear and pointwise."""
    rng = np.random.default_rng(14)
    s = rng.standard_normal(1000)
    z = np.zeros_like(s)
    assert np.allclose(O.pressure(s, s), -s, rtol=1e-15, atol=0)
    assert np.allclose(O.pressure(s, z), -2.0 / 3.0 * s, rtol=1e-15, atol=0)
    assert np.allclose(O.pressure(z, s), -1.0 / 3.0 * s, rtol=1e-15, atol=0)
    a, b = rng.standard_normal(1000), rng.standard_normal(1000)
    assert np.allclose(O.pressure(a, b), O.pressure(a, z) + O.pressure(z, b), rtol=1e-14, atol=1e-15)
    # pointwise: permuting the points permutes the pressures
    perm = rng.permutation(1000)
    assert np.array_equal(O.pressure(a[perm], b[perm]), O.pressure(a, b)[perm])
