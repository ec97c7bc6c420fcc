"""Tensor-level wrappers of the C ABI (PyTorch supplies device memory and the stream; it is plumbing only).

Every function here launches the gfx950 kernels through libtti_hip.so on torch's current stream. Volumes are
contiguous float32 CUDA tensors of shape grid.shape; sponge arrays are float32 tensors of shape (n, 2) = {A, B}.
Names follow the reference operators (kernels/kernel_type2.h, kernels/kernel_base.h).
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import call
from .grid import Grid


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _ptr(t: torch.Tensor | None, grid: Grid | None = None, dtype=torch.float32) -> int:
    if t is None:
        return 0
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous():
        raise ValueError("expected a contiguous CUDA tensor of dtype %s" % dtype)
    if grid is not None and tuple(t.shape) != grid.shape:
        raise ValueError(f"tensor shape {tuple(t.shape)} does not match grid {grid.shape}")
    return t.data_ptr()


def fd_consts(dx: float, dy: float, dz: float):
    """kernel_utils::compute_fd_const (kernels/kernel_utils.h:83-97): computed in double, stored as float."""
    f = lambda v: float(torch.tensor(v, dtype=torch.float32))
    return dict(
        invdx=f(1.0 / dx), invdy=f(1.0 / dy), invdz=f(1.0 / dz),
        invdxx=f(1.0 / (dx * dx)), invdyy=f(1.0 / (dy * dy)), invdzz=f(1.0 / (dz * dz)),
        invdxy=f(1.0 / (4 * dx * dy)), invdyz=f(1.0 / (4 * dy * dz)), invdzx=f(1.0 / (4 * dx * dz)),
    )


def fwd_drv(g: Grid, f1, f2, b, d, dx, dy, dz, order: int = 8):
    """d = [d1x, d1y, d1z, d2x, d2y, d2z] = b * D- (f1, f2)  (kernel_type2::fwd_main_loop)."""
    c = fd_consts(dx, dy, dz)
    call("launch_fwd_kernel2_drv", _ptr(f1, g), _ptr(f2, g), _ptr(b, g), *[_ptr(x, g) for x in d], *g.ranges,
         c["invdx"], c["invdy"], c["invdz"], g.ldx, g.ldy, order, _stream())


def fwd_update(g: Grid, u, w, d, f1, f2, models, dt, dx, dy, dz, order: int = 8):
    """u, w += dt * L(f, d)  (kernel_type2::fwd_main_loop_2). models = [m1, m2, m3, m4, rx, ry, rz, b]."""
    m1, m2, m3, m4, rx, ry, rz, b = models
    d1x, d1y, d1z, d2x, d2y, d2z = d
    call("launch_update_rho_fwd_main_loop_2_1", _ptr(u, g), _ptr(w, g), _ptr(d1x, g), _ptr(d2x, g), _ptr(d1y, g),
         _ptr(d2y, g), _ptr(d1z, g), _ptr(d2z, g), _ptr(f1, g), _ptr(f2, g), _ptr(b, g), _ptr(m1, g), _ptr(m2, g),
         _ptr(m3, g), _ptr(m4, g), _ptr(rx, g), _ptr(ry, g), _ptr(rz, g), *([0] * 12), *g.ranges, g.ldx, g.ldy,
         float(dt), float(dx), float(dy), float(dz), 0, order, 0, 0, 0, _stream())


def adj_drv(g: Grid, d, rhs3, rhs4, f1, f2, models, dx, dy, dz, order: int = 8):
    """Adjoint derivatives (kernel_type2::adj_main_loop): d = b D- S, rhs = sum X_ab M_ab."""
    m1, m2, m3, m4, rx, ry, rz, b = models
    d1x, d1y, d1z, d2x, d2y, d2z = d
    c = fd_consts(dx, dy, dz)
    call("launch_adj_kernel2_drv", _ptr(d1x, g), _ptr(d2x, g), _ptr(d1y, g), _ptr(d2y, g), _ptr(d1z, g),
         _ptr(d2z, g), _ptr(rhs3, g), _ptr(rhs4, g), _ptr(f1, g), _ptr(f2, g), _ptr(b, g), _ptr(m1, g), _ptr(m2, g),
         _ptr(m3, g), _ptr(m4, g), _ptr(rx, g), _ptr(ry, g), _ptr(rz, g), *g.ranges, c["invdxx"], c["invdyy"],
         c["invdzz"], c["invdxy"], c["invdyz"], c["invdzx"], g.ldx, g.ldy, order, 0.0, c["invdx"], c["invdy"],
         c["invdz"], _stream())


def adj_update(g: Grid, u, w, d, rhs3, rhs4, dt, dx, dy, dz, sponge=None, order: int = 8):
    """u, w += dt (D+ d + rhs)  or the sponge form (kernel_type2::adj_main_loop_2). sponge = (abx, aby, abz)."""
    d1x, d1y, d1z, d2x, d2y, d2z = d
    abx, aby, abz = sponge if sponge is not None else (None, None, None)
    call("launch_update_adj_kernel_main_3", _ptr(u, g), _ptr(w, g), _ptr(d1x, g), _ptr(d2x, g), _ptr(d1y, g),
         _ptr(d2y, g), _ptr(d1z, g), _ptr(d2z, g), _ptr(rhs3, g), _ptr(rhs4, g), _ptr(abx), _ptr(aby), _ptr(abz),
         *g.ranges, g.ldx, g.ldy, float(dt), float(dx), float(dy), float(dz), 0, 0, 0, order,
         1 if sponge is not None else 0, _stream())


def rho_gradient(g: Grid, grad, fwd_d, f1, f2, models, dx, dy, dz, grad_ldx: int, grad_ldy: int, order: int = 8):
    """grad += -b sum[S D+ Dfwd + (D- S) Dfwd] (kernel_type2::update_grad_loop_2). grad is a no-shift buffer."""
    m1, m2, m3, m4, rx, ry, rz, b = models
    f1x, f1y, f1z, f2x, f2y, f2z = fwd_d
    call("launch_update_rho_gradient_rho", _ptr(grad), _ptr(f1x, g), _ptr(f2x, g), _ptr(f1y, g), _ptr(f2y, g),
         _ptr(f1z, g), _ptr(f2z, g), _ptr(f1, g), _ptr(f2, g), _ptr(b, g), _ptr(m1, g), _ptr(m2, g), _ptr(m3, g),
         _ptr(m4, g), _ptr(rx, g), _ptr(ry, g), _ptr(rz, g), *g.ranges, g.ldx, g.ldy, grad_ldx, grad_ldy, float(dx),
         float(dy), float(dz), order, 0, _stream())


def loop2(g: Grid, f1, f2, u, w, abx, aby, abz, dt):
    """f = A f + B dt u (kernel_base::kernel_loop2)."""
    call("kernel_loop2_inner", _ptr(f1, g), _ptr(f2, g), _ptr(u, g), _ptr(w, g), _ptr(abx), _ptr(aby), _ptr(abz),
         *g.ranges, g.ldx, g.ldy, float(dt), _stream())


def _pvec(ts, g: Grid):
    return (C.c_void_p * len(ts))(*[_ptr(t, g) for t in ts])


def _fvec(vals):
    return (C.c_float * len(vals))(*[float(x) for x in vals])


def visco_loop1(g: Grid, f1, f2, u, w, m8, m9, wsi, abx, aby, abz, dt):
    """f = A f + B dt (u - sum_m wsi[m] M_m) with memory variables m8 (of f1) and m9 (of f2) (kernel_base::kernel_loop1);
    abx = aby = abz = None for A = B = 1."""
    call("kernel_loop1_inner", _ptr(f1, g), _ptr(f2, g), _ptr(u, g), _ptr(w, g), _pvec(m8, g), _pvec(m9, g),
         _fvec(wsi), len(wsi), _ptr(abx), _ptr(aby), _ptr(abz), *g.ranges, g.ldx, g.ldy, float(dt), _stream())


def visco_loop3(g: Grid, u, w, m8, m9, m10, wsi, dt):
    """M_m += dt (alpha_m u - wsi[m] M_m), alpha = m10 (kernel_base::kernel_loop3)."""
    call("kernel_loop3_inner", _ptr(u, g), _ptr(w, g), _pvec(m8, g), _pvec(m9, g), _pvec(m10, g), _fvec(wsi),
         len(wsi), *g.ranges, g.ldx, g.ldy, float(dt), _stream())


def visco_loop3_born(g: Grid, u, w, u1, w1, m8, m9, m81, m91, m10, wsi, dt):
    """visco_loop3 on (u, w, m8, m9) and on the Born set (u1, w1, m81, m91) in one pass with the same alpha = m10
    (kernel_base::kernel_loop3 with update_born)."""
    call("kernel_loop3_inner_born", _ptr(u, g), _ptr(w, g), _ptr(u1, g), _ptr(w1, g), _pvec(m8, g), _pvec(m9, g),
         _pvec(m81, g), _pvec(m91, g), _pvec(m10, g), _fvec(wsi), len(wsi), *g.ranges, g.ldx, g.ldy, float(dt),
         _stream())


def visco_adj_loop1(g: Grid, f1, f2, u, w, m8, m9, m10, abx, aby, abz, dt):
    """f = A f + B dt (u + sum_m alpha_m M_m) (kernel_base::adj_kernel_loop1)."""
    call("adj_kernel_loop1_inner", _ptr(f1, g), _ptr(f2, g), _ptr(u, g), _ptr(w, g), _pvec(m8, g), _pvec(m9, g),
         _pvec(m10, g), len(m8), _ptr(abx), _ptr(aby), _ptr(abz), *g.ranges, g.ldx, g.ldy, float(dt), _stream())


def visco_adj_loop3(g: Grid, u, w, m8, m9, wsi, dt):
    """M_m += dt wsi[m] (-u - M_m) (kernel_base::adj_kernel_loop3)."""
    call("adj_kernel_loop3_inner", _ptr(u, g), _ptr(w, g), _pvec(m8, g), _pvec(m9, g), _fvec(wsi), len(wsi),
         *g.ranges, g.ldx, g.ldy, float(dt), _stream())


def loop4(g: Grid, f1, f2):
    """Free-surface mirror of sigma at the top interior plane (kernel_base::kernel_loop4)."""
    call("kernel_loop4_inner", _ptr(f1, g), _ptr(f2, g), g.ldx, g.ldy, g.ghost, _stream())


def loop4_derivatives(g: Grid, d):
    """Free-surface mirror of the derivatives (kernel_type2::kernel_loop4_derivative)."""
    call("kernel_loop4_derivatives", *[_ptr(x, g) for x in d], g.ldx, g.ldy, g.ghost, _stream())


def vp_gradient(g: Grid, snap3, snap4, grad, adj3, adj4):
    """grad += adj3*snap3 + adj4*snap4 with snap/grad of interior shape (nz, ny, nx)."""
    call("update_Vp_gradient", _ptr(snap3), _ptr(snap4), _ptr(grad), _ptr(adj3, g), _ptr(adj4, g), *g.ranges,
         g.nx, g.ny, g.ldx, g.ldy, _stream())


def finalize_vp_gradient(g: Grid, model3, grad, dt_xcorr):
    call("finalize_Vp_gradient", _ptr(model3, g), _ptr(grad), *g.ranges, g.nx, g.ny, g.ldx, g.ldy,
         float(dt_xcorr), _stream())


def init_cijs(vp, eps, delta, vs, theta, phi, rho, out):
    """out = [m1, m2, m3, m4, rx, ry, rz, b] from per-point physical parameters (all of the same numel)."""
    n = vp.numel()
    call("init_cijs", n, *[_ptr(t) for t in (vp, eps, delta, vs, theta, phi, rho)], *[_ptr(t) for t in out],
         _stream())


def pressure(f1, f2, p):
    call("compute_pressure", f1.numel(), _ptr(f1), _ptr(f2), _ptr(p), _stream())


def t1_fwd(g: Grid, u, w, f1, f2, models7, dt, dx, dy, dz, mode: int = 0, sponge=None, out_ld=(0, 0), order: int = 8):
    """Constant-density TTI forward (kernel_type1_gpu::fwd_main_loop_2): mode 0 u += dt rhs, 1 sponge form,
    2 calc_snap (u, w receive rhs in a no-halo layout with leading dims out_ld)."""
    m = [_ptr(t, g) for t in models7]
    abx, aby, abz = sponge if sponge is not None else (None, None, None)
    call("fwd_main_loop_2_1_inner_gpu_launch", _ptr(u, None if mode == 2 else g), _ptr(w, None if mode == 2 else g),
         _ptr(f1, g), _ptr(f2, g), *m, _ptr(abx), _ptr(aby), _ptr(abz), *g.ranges, g.ldx, g.ldy, g.ldz,
         int(out_ld[0]), int(out_ld[1]), 0, float(dt), float(dx), float(dy), float(dz), order, _stream(),
         1 if mode == 2 else 0, 1 if mode == 1 else 0, 0)


def t1_adj(g: Grid, u, w, f1, f2, models7, dt, dx, dy, dz, sponge=None, order: int = 8):
    """Constant-density TTI adjoint (kernel_type1_gpu::adj_main_loop_2)."""
    m = [_ptr(t, g) for t in models7]
    abx, aby, abz = sponge if sponge is not None else (None, None, None)
    call("adj_main_loop_2_inner_gpu_launch", _ptr(u, g), _ptr(w, g), _ptr(f1, g), _ptr(f2, g), *m, _ptr(abx),
         _ptr(aby), _ptr(abz), *g.ranges, g.ldx, g.ldy, g.ldz, float(dt), float(dx), float(dy), float(dz), order,
         _stream(), 1 if sponge is not None else 0, 0)


# ---- eTTI operators of the simulator3 driver (include/tti_etti.h) ----------------------------------------------------

def etti_set_coefficients(df=None, si=None):
    """Coefficient tables of the eTTI stencils (etti_init_constant); None keeps a table. Both start at zero."""
    f4 = lambda v: None if v is None else (C.c_float * 4)(*[float(x) for x in v])
    call("etti_init_constant", f4(df), f4(si))


def etti_coefficients():
    df, si = (C.c_float * 4)(), (C.c_float * 4)()
    call("etti_get_constant", df, si)
    return list(df), list(si)


def _ms(ms):
    return None if ms is None else C.byref(ms)


def etti_velocity(g: Grid, v, s, M22, zprime, ms=None):
    """v = [f1, f2, f3] += ... from stresses s = [f4..f9] (vPlainLoop). ms: optional ctypes.c_float accumulator."""
    call("vPlainLoop", *[_ptr(t, g) for t in v], *[_ptr(t, g) for t in s], _ptr(M22, g), _ptr(zprime), g.nx, g.ny,
         g.nz, g.ldx, g.ldy, _ms(ms), _stream())


def etti_strain(g: Grid, v, d, zprime, ms=None):
    """d = [d4..d9] from velocities v = [f1, f2, f3] (vgPlainLoop)."""
    call("vgPlainLoop", *[_ptr(t, g) for t in v], *[_ptr(t, g) for t in d], _ptr(zprime), g.nx, g.ny, g.nz, g.ldx,
         g.ldy, _ms(ms), _stream())


def etti_mixed(g: Grid, src, xz, yz, xs, ys, zs):
    """xz, yz = x-z / y-z diagonal interpolations of src with shifts (precompute_xz_yz)."""
    call("precompute_xz_yz", int(xs), int(ys), int(zs), _ptr(src, g), _ptr(xz, g), _ptr(yz, g), g.nx, g.ny, g.nz,
         g.ldx, g.ldy, _stream())


def etti_stress_loop(g: Grid, s, M, d, x, ms=None):
    """sPlainLoop: s = [f4..f9] updated, M = [M1..M21], d = [d4..d9], x = [f10..f21] (written by the mixes)."""
    call("sPlainLoop", *[_ptr(t, g) for t in s], *[_ptr(t, g) for t in M], *[_ptr(t, g) for t in d],
         *[_ptr(t, g) for t in x], g.nx, g.ny, g.nz, g.ldx, g.ldy, _ms(ms), _stream())


def etti_stress(g: Grid, s, M, d, x):
    """The stress update alone (sPlainKernel_launch), with x = [f10..f21] as inputs."""
    arr = lambda ts: (C.c_void_p * len(ts))(*[_ptr(t, g) for t in ts])
    call("sPlainKernel_launch", arr(s), arr(M), arr(d), arr(x), g.nx, g.ny, g.nz, g.ldx, g.ldy, _stream())


def etti_mirror(g: Grid, fields, izbeg=None):
    """Free surface over whole x-y planes 0..izbeg (mirror_s_inner for 6 fields, mirror_v_inner for 3)."""
    z = g.ghost if izbeg is None else izbeg
    name = {6: "mirror_s_inner", 3: "mirror_v_inner"}[len(fields)]
    call(name, *[_ptr(t, g) for t in fields], g.ldx * g.ldy, z, _stream())


def etti_update_gv(g: Grid, snap1, snap3, g1, a1, a3):
    """g1 += a1 snap1 + a3 snap3 (update_gv_inner); g1 and the snapshots have the interior shape (nz, ny, nx)."""
    call("update_gv_inner", _ptr(snap1), _ptr(snap3), _ptr(g1), _ptr(a1, g), _ptr(a3, g), *g.ranges, g.nx, g.ny,
         g.ldx, g.ldy, _stream())


def etti_init_cijs(g: Grid, M, d, zprime):
    """M = [M1..M22], d = [d4..d9] volumes and zprime (length ldz) of the synthetic model (etti_init_cijs)."""
    arr = lambda ts: (C.c_void_p * len(ts))(*[_ptr(t, g) for t in ts])
    call("etti_init_cijs", arr(M), arr(d), _ptr(zprime), g.ghost, g.ghost + g.nx - 1, g.ldx, g.ldy, g.ldz,
         _stream())


def etti_pressure(p, s):
    call("etti_compute_p", _ptr(p), *[_ptr(t) for t in s], p.numel(), _stream())
