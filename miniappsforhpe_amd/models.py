"""Synthetic TTI earth models, source wavelet and time-step rule (generated on the GPU with torch; plumbing only).

The stiffness-like model volumes follow the reference's init_cijs (simulator2/simulator2SetUp.cu:10-29):
  m1 = vp^2 (1 + 2 eps), m2 = sqrt((vp^2 - vs^2)(vp^2 (1 + 2 delta) - vs^2)), m3 = vp^2, m4 = vs^2,
  (rx, ry, rz) = (sin t cos p, sin t sin p, cos t), b = 1/rho.
Unlike the reference (which leaves Model11 = 1/rho unset and the x ghosts of the models at zero), every model volume
is filled everywhere, ghosts included, by edge replication of the interior.
"""
from __future__ import annotations

import math

import torch

from .grid import Grid


def gaussian_smooth(v: torch.Tensor, sigma: float) -> torch.Tensor:
    """Separable Gaussian smoothing (sigma in cells) of a 3-D tensor with replicate padding (shift-and-add, so no
    convolution library is involved)."""
    if sigma <= 0:
        return v
    r = int(math.ceil(3 * sigma))
    x = torch.arange(-r, r + 1, dtype=torch.float64)
    k = torch.exp(-0.5 * (x / sigma) ** 2)
    k = (k / k.sum()).tolist()
    out = v
    for axis in range(3):
        n = out.shape[axis]
        first = out.narrow(axis, 0, 1)
        last = out.narrow(axis, n - 1, 1)
        padded = torch.cat([first.expand(*[r if a == axis else out.shape[a] for a in range(3)]), out,
                            last.expand(*[r if a == axis else out.shape[a] for a in range(3)])], dim=axis)
        acc = torch.zeros_like(out)
        for j in range(2 * r + 1):
            acc += k[j] * padded.narrow(axis, j, n)
        out = acc
    return out


def fill_ghosts(grid: Grid, interior: torch.Tensor) -> torch.Tensor:
    """Padded volume whose interior is `interior` and whose ghosts/padding replicate the nearest interior value."""
    g = grid.ghost
    v = torch.nn.functional.pad(interior[None, None], (g, grid.ldx - grid.nx - g, g, g, g, g), mode="replicate")
    return v[0, 0].contiguous()


def layered_tti(grid: Grid, seed: int, device="cuda", smooth: float = 8.0, vmin: float = 1500.0,
                vmax: float = 4500.0, layers: int = 8, noise: float = 0.05, vs_ratio: float = 0.0,
                rho: float = 2000.0):
    """Layered vp (vmin..vmax in `layers` z layers) with seeded Gaussian-smoothed +-noise, eps ~ U(0, 0.3),
    delta ~ U(0, 0.15), tilt ~ U(0, 60 deg), azimuth ~ U(0, 360 deg), each smoothed with `smooth` cells.
    Returns (models [m1, m2, m3, m4, rx, ry, rz, b], vmax_effective)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    shp = (grid.nz, grid.ny, grid.nx)
    zi = torch.arange(grid.nz, device=device, dtype=torch.float32)
    layer = torch.clamp((zi * layers / grid.nz).floor(), max=layers - 1)
    v1d = vmin + (vmax - vmin) * layer / max(layers - 1, 1)

    def field(lo, hi):
        u = torch.rand(shp, generator=gen, device=device) * (hi - lo) + lo
        return gaussian_smooth(u, smooth)

    def noise_field():
        n = torch.randn(shp, generator=gen, device=device)
        n = gaussian_smooth(n, smooth)
        return n / n.abs().amax().clamp_min(1e-30)

    vp = v1d[:, None, None] * (1 + noise * noise_field())
    eps = field(0.0, 0.3)
    delta = field(0.0, 0.15)
    theta = field(0.0, math.pi / 3)
    phi = field(0.0, 2 * math.pi)
    vs = vs_ratio * vp
    dens = torch.full(shp, rho, device=device)
    return build_models(grid, vp, eps, delta, vs, theta, phi, dens)


def build_models(grid: Grid, vp, eps, delta, vs, theta, phi, rho):
    """Model volumes from interior-shaped physical parameters (FP32), ghosts replicated."""
    v2, s2 = vp * vp, vs * vs
    m = [v2 * (1 + 2 * eps), torch.sqrt((v2 - s2) * (v2 * (1 + 2 * delta) - s2)), v2, s2,
         torch.sin(theta) * torch.cos(phi), torch.sin(theta) * torch.sin(phi), torch.cos(theta), 1.0 / rho]
    vmax = float(torch.sqrt(m[0].amax()))
    return [fill_ghosts(grid, t.float()) for t in m], vmax


def cfl_dt(vmax: float, dx: float, dy: float, dz: float, fraction: float = 0.91) -> float:
    """Reference dt rule: 0.77 * fraction / (v sqrt(1/dx^2 + 1/dy^2 + 1/dz^2)) (simulator2/simulator2SetUp.cu:127-129),
    with v the fastest phase velocity vp sqrt(1 + 2 eps)."""
    return fraction * 0.77 / (vmax * math.sqrt(1 / dx**2 + 1 / dy**2 + 1 / dz**2))


def ricker(nt: int, dt: float, f0: float, t0: float | None = None) -> torch.Tensor:
    """Ricker wavelet (1 - 2 pi^2 f0^2 tau^2) exp(-pi^2 f0^2 tau^2), tau = t - t0, t0 = 1.2/f0 by default."""
    t0 = 1.2 / f0 if t0 is None else t0
    t = torch.arange(nt, dtype=torch.float64) * dt - t0
    a = (math.pi * f0 * t) ** 2
    return ((1 - 2 * a) * torch.exp(-a)).float()


def cerjan_sponge(n_with_ghost: int, ghost: int, n: int, width: int, alpha: float = 0.015):
    """Cerjan et al. (1985) taper A(i) = exp(-(alpha (width - i))^2) over `width` interior points at each end,
    A = B. Returns a (n_with_ghost, 2) float32 tensor {A, B} indexed by array index."""
    ab = torch.ones((n_with_ghost, 2), dtype=torch.float32)
    for i in range(width):
        a = math.exp(-((alpha * (width - i)) ** 2))
        ab[ghost + i] = a
        ab[ghost + n - 1 - i] = a
    return ab
