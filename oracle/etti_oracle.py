"""TEST INFRASTRUCTURE ONLY: FP64 numpy restatement of the simulator3 eTTI (staggered elastic) operators.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and only as the
reference that GPU results are compared against. The product package never imports it.

Volumes are float64 arrays of shape (ldz, ldy, ldx) with 4-point ghosts; every operator writes the interior
[4, 4+n) of each axis only, as the reference kernels do when n is a multiple of their tile.

Coefficients. The reference keeps them in constant memory (simulator3/simulator3SetUp.cu:25-26), filled by
init_constant (:28-34) with {1.01, 2.01, 3.01, 4.01} for both tables; no code path calls init_constant, so the
reference runs with zeros. Every function here takes the two 4-entry tables explicitly:
  df -- staggered first derivative: D+ g(i) = sum_k df[k] (g(i+k+1) - g(i-k)),  D- g(i) = sum_k df[k] (g(i+k) - g(i-k-1))
        (drv1_8<SHIFT=false/true>, :36-45)
  si -- diagonal four-corner interpolation between two axes a, b with shifts (sa, sb):
        Mix g = 0.5 sum_{n=1..4} si[n-1] sum_{p in {n, 1-n}} sum_{q in {n, 1-n}} g(i_a + p - sa, i_b + q - sb)
        (mixed_XY, :207-220; precompute_xz_yz main loop, :137-175)
The velocity kernel scales with invd1 = 1.01, invd2 = 2.01 and the strain kernel with invd1 = 1.01, invd2 = 2.02,
both with zprime[z] on z terms (:833-834, :1240-1241); the defaults below are those constants.
"""
from __future__ import annotations

import numpy as np

G = 4  # ghost width

V_INVD = (1.01, 2.01)   # vPlainKernel (simulator3/simulator3SetUp.cu:833-834)
VG_INVD = (1.01, 2.02)  # vgPlainKernel (:1240-1241)
REF_CONSTANT = (1.01, 2.01, 3.01, 4.01)  # init_constant (:30-31), never uploaded by the reference


def win(a, nx, dz=0, dy=0, dx=0):
    """Interior window of `a` (x extent nx) shifted by (dz, dy, dx)."""
    nz, ny = a.shape[0] - 2 * G, a.shape[1] - 2 * G
    return a[G + dz: G + dz + nz, G + dy: G + dy + ny, G + dx: G + dx + nx]


_AX = {"x": 2, "y": 1, "z": 0}


def _sh(axis, k):
    d = [0, 0, 0]
    d[_AX[axis]] = k
    return d


def dplus(a, nx, axis, df):
    """D+ along `axis` on the interior (value at i + 1/2)."""
    return sum(df[k] * (win(a, nx, *_sh(axis, k + 1)) - win(a, nx, *_sh(axis, -k))) for k in range(4))


def dminus(a, nx, axis, df):
    """D- along `axis` on the interior (value at i - 1/2)."""
    return sum(df[k] * (win(a, nx, *_sh(axis, k)) - win(a, nx, *_sh(axis, -k - 1))) for k in range(4))


def mix(a, nx, axes, shifts, si):
    """Diagonal four-corner interpolation in the plane of `axes` (two of 'x', 'y', 'z') with shifts."""
    (ea, eb), (sa, sb) = axes, shifts
    out = 0.0
    for n in range(1, 5):
        for p in (n, 1 - n):
            for q in (n, 1 - n):
                d = [0, 0, 0]
                d[_AX[ea]] += p - sa
                d[_AX[eb]] += q - sb
                out = out + si[n - 1] * win(a, nx, *d)
    return 0.5 * out


def velocity(nx, f1, f2, f3, f4, f5, f6, f7, f8, f9, M22, zprime, df, invd=V_INVD):
    """f1..f3 += 2 M22(shifted) (...) -- vPlainKernel (simulator3/simulator3SetUp.cu:820-975)."""
    i1, i2 = invd
    zp = zprime[G: f1.shape[0] - G][:, None, None]
    r1 = 2.0 * win(M22, nx, 0, 0, -1) * (dminus(f4, nx, "x", df) * i1 + dplus(f7, nx, "y", df) * i2
                                        + dplus(f8, nx, "z", df) * zp)
    r2 = 2.0 * win(M22, nx, 0, -1, 0) * (dplus(f7, nx, "x", df) * i1 + dminus(f5, nx, "y", df) * i2
                                        + dplus(f9, nx, "z", df) * zp)
    r3 = 2.0 * win(M22, nx, -1, 0, 0) * (dplus(f8, nx, "x", df) * i1 + dplus(f9, nx, "y", df) * i2
                                        + dminus(f6, nx, "z", df) * zp)
    win(f1, nx)[...] += r1
    win(f2, nx)[...] += r2
    win(f3, nx)[...] += r3


def strain(nx, f1, f2, f3, d4, d5, d6, d7, d8, d9, zprime, df, invd=VG_INVD):
    """d4..d9 from the velocities -- vgPlainKernel (simulator3/simulator3SetUp.cu:1218-1361)."""
    i1, i2 = invd
    zp = zprime[G: f1.shape[0] - G][:, None, None]
    win(d4, nx)[...] = i1 * dplus(f1, nx, "x", df)
    win(d5, nx)[...] = i2 * dminus(f1, nx, "y", df) + i1 * dminus(f2, nx, "x", df)
    win(d6, nx)[...] = zp * dminus(f1, nx, "z", df) + i1 * dminus(f3, nx, "x", df)
    win(d7, nx)[...] = i2 * dplus(f2, nx, "y", df)
    win(d8, nx)[...] = zp * dminus(f2, nx, "z", df) + i2 * dminus(f3, nx, "y", df)
    win(d9, nx)[...] = zp * dplus(f3, nx, "z", df)


def mixed_xz_yz(nx, g, xz, yz, xs, ys, zs, si):
    """xz = Mix_xz(g; xs, zs), yz = Mix_yz(g; ys, zs) -- precompute_xz_yz<XS,YS,ZS> (:47-176), with the coefficient
    of each plane taken by its offset from the output plane throughout (the reference's priming loop, :120-125,
    indexes it by the priming step for the first outputs of a z block; with its zero coefficients no output shows
    the difference)."""
    win(xz, nx)[...] = mix(g, nx, ("x", "z"), (xs, zs), si)
    win(yz, nx)[...] = mix(g, nx, ("y", "z"), (ys, zs), si)


# (input strain, xz output, yz output, XS, YS, ZS) of the six precompute launches in sPlainLoop (:1186-1191)
PRECOMPUTE = (
    ("d4", "f10", "f11", 1, 1, 1),
    ("d7", "f16", "f17", 1, 1, 1),
    ("d9", "f20", "f21", 1, 1, 1),
    ("d5", "f12", "f13", 0, 0, 1),
    ("d6", "f14", "f15", 0, 1, 0),
    ("d8", "f18", "f19", 1, 0, 0),
)


def stress(nx, f, M, d, si):
    """f4..f9 += C(M1..M21) . strains -- sPlainKernel (simulator3/simulator3SetUp.cu:977-1104).

    f: dict with f4..f9 (updated) and f10..f21 (read); M: list M[1]..M[21] (M[0] unused); d: dict d4..d9."""
    w = lambda a: win(a, nx)
    m = [None] + [w(M[i]) for i in range(1, 22)]
    dv1d2_xx = mix(d["d5"], nx, ("x", "y"), (0, 0), si)
    dv1d1_xy = mix(d["d4"], nx, ("x", "y"), (1, 1), si)
    dv2d2_xy = mix(d["d7"], nx, ("x", "y"), (1, 1), si)
    dv3d3_xy = mix(d["d9"], nx, ("x", "y"), (1, 1), si)
    dv2d3_xz = mix(d["d8"], nx, ("x", "y"), (1, 0), si)
    dv1d3_yz = mix(d["d6"], nx, ("x", "y"), (0, 1), si)
    d4, d5, d6, d7, d8, d9 = (w(d[k]) for k in ("d4", "d5", "d6", "d7", "d8", "d9"))
    F = {k: w(f[k]) for k in f}
    F["f4"][...] += m[1] * d4 + m[2] * d7 + m[3] * d9 + m[6] * dv1d2_xx + m[5] * F["f14"] + m[4] * F["f19"]
    F["f5"][...] += m[2] * d4 + m[7] * d7 + m[8] * d9 + m[11] * dv1d2_xx + m[10] * F["f14"] + m[9] * F["f19"]
    F["f6"][...] += m[3] * d4 + m[8] * d7 + m[12] * d9 + m[15] * dv1d2_xx + m[14] * F["f14"] + m[13] * F["f19"]
    F["f7"][...] += m[6] * dv1d1_xy + m[11] * dv2d2_xy + m[15] * dv3d3_xy + m[20] * F["f15"] + m[18] * F["f18"] \
        + m[21] * d5
    F["f8"][...] += m[5] * F["f10"] + m[10] * F["f16"] + m[14] * F["f20"] + m[20] * F["f13"] + m[17] * dv2d3_xz \
        + m[19] * d6
    F["f9"][...] += m[4] * F["f11"] + m[9] * F["f17"] + m[13] * F["f21"] + m[18] * F["f12"] + m[17] * dv1d3_yz \
        + m[16] * d8


def stress_loop(nx, f, M, d, si):
    """sPlainLoop (:1106-1216): the six precompute launches, then the stress update."""
    for src, oxz, oyz, xs, ys, zs in PRECOMPUTE:
        mixed_xz_yz(nx, d[src], f[oxz], f[oyz], xs, ys, zs, si)
    stress(nx, f, M, d, si)


def mirror(fields, izbeg=G):
    """Free surface at plane izbeg over the whole x-y plane: zero at izbeg, odd mirror above
    (mirror_s_kernel / mirror_v_kernel, simulator3/simulator3SetUp.cu:1379-1480)."""
    for a in fields:
        a[izbeg] = 0.0
        for iz in range(izbeg):
            a[iz] = -a[2 * izbeg - iz]


def update_gv(nx, g1, snap1, snap3, a1, a3):
    """g1 += a1 snap1 + a3 snap3 on the interior; g1 and the snapshots have no ghosts (update_gv, :1482-1500)."""
    g1 += win(a1, nx) * snap1 + win(a3, nx) * snap3


def compute_p(f4, f5, f6, f7, f8, f9):
    """compute_p_kernel (:273-287) over the whole allocation."""
    return -(2 * f4 + 3 * f5 + 4 * f6 + 5 * f7 + 6 * f8 + 7 * f9) / 27.0


# init_cijs_kernel (:230-271): volume -> (cos|sin, weights of (iz, iy, ix)) for x in [ixbeg, ixend], all y and z
CIJS = [("M%d" % (i + 1), f, w) for i, (f, w) in enumerate([
    (np.cos, (2, 1, 1)), (np.sin, (2, 1, 1)), (np.cos, (1, 2, 1)), (np.sin, (1, 2, 1)), (np.cos, (1, 1, 2)),
    (np.sin, (1, 1, 2)), (np.cos, (3, 1, 1)), (np.sin, (3, 1, 1)), (np.cos, (1, 3, 1)), (np.sin, (1, 3, 1)),
    (np.cos, (1, 1, 3)), (np.sin, (1, 1, 3)), (np.cos, (4, 1, 1)), (np.sin, (4, 1, 1)), (np.cos, (1, 4, 1)),
    (np.sin, (1, 4, 1)), (np.cos, (1, 1, 4)), (np.sin, (1, 1, 4)), (np.cos, (5, 1, 1)), (np.sin, (5, 1, 1)),
    (np.cos, (1, 5, 1)), (np.sin, (1, 5, 1))])] + [
    ("d4", np.cos, (1, 1, 5)), ("d5", np.sin, (1, 1, 5)), ("d6", np.cos, (6, 1, 1)), ("d7", np.sin, (6, 1, 1)),
    ("d8", np.cos, (1, 6, 1)), ("d9", np.sin, (1, 6, 1))]


def init_cijs(shape, nx):
    """The 28 model/strain volumes and zprime of init_cijs_kernel; the argument is formed in FP32 as the kernel does
    (w_z*iz + w_y*iy + w_x*ix with float weights), the cos/sin in FP64."""
    ldz, ldy, ldx = shape
    iz, iy, ix = np.meshgrid(np.arange(ldz), np.arange(ldy), np.arange(G, G + nx), indexing="ij")
    out = {}
    for name, fn, (wz, wy, wx) in CIJS:
        arg = (np.float32(wz) * iz.astype(np.float32) + np.float32(wy) * iy.astype(np.float32)).astype(np.float32)
        arg = (arg + np.float32(wx) * ix.astype(np.float32)).astype(np.float32)
        v = np.zeros(shape)
        v[:, :, G: G + nx] = fn(arg.astype(np.float64))
        out[name] = v
    zprime = np.cos((np.float32(6.0) * np.arange(ldz, dtype=np.float32)).astype(np.float64))
    return out, zprime


# ---- the simulator3 driver over sub-volumes (simulator3/simulator3SetUp.cu:2181-2657) ------------------------------

class Sim3:
    """FP64 replay of the simulator3 step sequence over per-sub-volume arrays with the X -> Y -> Z ghost exchange.

    subs: list of dicts {"offset": (ox, oy, oz), "n": (nx, ny, nz), "vols": {name: float64 array}, "zprime": array}
    with every volume of the driver (f1..f21, M1..M22, d4..d9, g2) of shape (nz+8, ny+8, ldx) and g1 of (nz, ny, nx).
    Free-surface mirrors act on the sub-volumes at the global top (z offset 0)."""

    S_FIELDS = ["f%d" % k for k in range(4, 10)]
    V_FIELDS = ["f1", "f2", "f3"]
    D_FIELDS = ["d%d" % k for k in range(4, 10)]

    def __init__(self, subs, df, si, free_surface=True):
        self.subs, self.df, self.si, self.free_surface = subs, df, si, free_surface
        self.at = {tuple(s["offset"]): s for s in subs}
        self.snaps = []

    def _nb(self, s, axis, side):
        o = list(s["offset"])
        if side > 0:
            o[axis] += s["n"][axis]
        else:
            cands = [t for t in self.subs if all(t["offset"][a] == o[a] for a in range(3) if a != axis)
                     and t["offset"][axis] + t["n"][axis] == o[axis]]
            return cands[0] if cands else None
        return self.at.get(tuple(o))

    def exchange(self, names):
        """Ghost fill X, then Y (x ghosts included), then Z (x and y ghosts included) from the neighbours' interiors;
        faces on the global boundary keep their ghosts (infrastructure/haloManager3D_gpu.cu:601-787)."""
        for axis in (0, 1, 2):  # x, y, z
            dim = 2 - axis
            for s in self.subs:
                nx, ny, nz = s["n"]
                ext = [slice(G, G + nz), slice(G, G + ny), slice(0, nx + 2 * G)]  # (z, y, x) extents of a face
                if axis == 0:
                    ext[2] = slice(G, G + nx)
                if axis == 2:
                    ext[1] = slice(0, ny + 2 * G)
                for side in (-1, 1):
                    t = self._nb(s, axis, side)
                    if t is None:
                        continue
                    n_s, n_t = s["n"][axis], t["n"][axis]
                    dst, src = list(ext), list(ext)
                    if side < 0:
                        dst[dim], src[dim] = slice(0, G), slice(n_t, n_t + G)
                    else:
                        dst[dim], src[dim] = slice(G + n_s, 2 * G + n_s), slice(G, 2 * G)
                    for name in names:
                        s["vols"][name][tuple(dst)] = t["vols"][name][tuple(src)]

    def _mirrors(self):
        if not self.free_surface:
            return
        for s in self.subs:
            if s["offset"][2] == 0:
                mirror([s["vols"][k] for k in self.S_FIELDS + self.V_FIELDS])

    def _velocity(self):
        for s in self.subs:
            v = s["vols"]
            velocity(s["n"][0], v["f1"], v["f2"], v["f3"], *[v["f%d" % k] for k in range(4, 10)], v["M22"],
                     s["zprime"], self.df)

    def _strain(self):
        for s in self.subs:
            v = s["vols"]
            strain(s["n"][0], v["f1"], v["f2"], v["f3"], *[v["d%d" % k] for k in range(4, 10)], s["zprime"], self.df)

    def _stress(self):
        for s in self.subs:
            v = s["vols"]
            f = {"f%d" % k: v["f%d" % k] for k in range(4, 22)}
            stress_loop(s["n"][0], f, [None] + [v["M%d" % k] for k in range(1, 22)],
                        {k: v[k] for k in self.D_FIELDS}, self.si)

    def source(self, total, p=500.0):
        c = [t // 2 for t in total]
        for s in self.subs:
            if all(s["offset"][a] <= c[a] < s["offset"][a] + s["n"][a] for a in range(3)):
                x, y, z = (c[a] - s["offset"][a] + G for a in range(3))
                for k in range(6):
                    s["vols"]["f%d" % (k + 4)][z, y, x] = p * (k + 1)

    def forward(self, niter, xcorr):
        for it in range(niter):
            self.exchange(self.S_FIELDS)
            self._velocity()
            self.exchange(self.V_FIELDS)
            self._mirrors()
            self._strain()
            self.exchange(self.D_FIELDS)
            self._stress()
            for s in self.subs:
                s["vols"]["p"] = compute_p(*[s["vols"][k] for k in self.S_FIELDS])
            if it % xcorr == 0:
                self.snaps.append([(win(s["vols"]["f1"], s["n"][0]).copy(), win(s["vols"]["f3"], s["n"][0]).copy())
                                   for s in self.subs])
                self.exchange(self.S_FIELDS)

    def adjoint(self, niter, xcorr):
        import oracle as O

        for it in range(niter - 1, -1, -1):
            self.exchange(self.V_FIELDS)
            self._stress()
            self.exchange(self.S_FIELDS)
            self._strain()
            self.exchange(self.D_FIELDS)
            self._mirrors()
            self._velocity()
            if it % xcorr:
                continue
            snap = self.snaps.pop()
            for s, (s1, s3) in zip(self.subs, snap):
                v = s["vols"]
                update_gv(s["n"][0], v["g1"], s1, s3, v["f1"], v["f3"])
            self.exchange(self.V_FIELDS)
            for s in self.subs:
                v = s["vols"]
                Gs = O.Geom(*s["n"])
                O.rho_gradient(Gs, v["g2"], [v["d%d" % k] for k in range(4, 10)], v["f4"], v["f6"],
                               [v["M%d" % k] for k in (1, 2, 3, 4, 5, 6, 7, 18)], 1.0, 1.0, 1.0, Gs.ldx, Gs.ldy)
