"""One-shot FWI gradient on the fused gfx950 time steps: forward with snapshots, receiver residual, adjoint with
residual injection and the zero-lag imaging condition accumulated in FP64.

Discrete system (no sponge; state before fused step n is (sigma_n, u_n)):
    u_{n+1}     = u_n + dt L(sigma_n) + s_n             (s_n = dt * wavelet[n] at the source, in u and w)
    sigma_{n+1} = sigma_n + B dt u_{n+1}                (B = 1)
    d_n         = P(sigma_{n+1}) at the receivers,      P = -(2 sigma_H + sigma_V) / 3
    J           = 1/2 sum_n |d_n - d_obs_n|^2
The adjoint of this recursion is the reference's adjoint step (kernels/kernel_type2_grad_snap.cu,
kernels/kernel_type2_adj.cu, kernels/kernel_base.cu:82-104) with the adjoint "u" holding mu_n, the multiplier of
sigma_n: after the adjoint step that yields mu_n, the residual of d_{n-1} is injected (P^T r: -2/3 r into u_H,
-1/3 r into u_V), and
    dJ/d epsilon(x) = dt * sum_n mu_n(x) u_n(x)          (B -> B (1 + epsilon(x)) in the sigma update)
which is the reference's imaging condition g += u_adj u_snap + w_adj w_snap (kernels/kernel_type2_small.cu:42-59)
before its -2 dt_xcorr / vp finalisation. With snapshots every k steps (n = k, 2k, ...) the sum is subsampled and
scaled by k, as the reference's dt_xcorr = k dt does.
"""
from __future__ import annotations

from dataclasses import dataclass

import torch

from ._lib import call
from .propagator import SnapshotStore, TTIPropagator


@dataclass
class Shot:
    src_index: int            # linear index of the source in the padded volume
    wavelet: torch.Tensor     # [nt] host float32; dt * wavelet[n] is added to u and w at step n
    rec_index: torch.Tensor   # [nrec] int64 linear indices (device)
    rec_ijk: torch.Tensor     # [3 * nrec] int32 array coordinates x, y, z of each receiver (device, distinct points)


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def forward(P: TTIPropagator, shot: Shot, nt: int, store: SnapshotStore | None = None, k: int = 1) -> torch.Tensor:
    """Runs nt fused forward steps from rest and returns the receiver pressures d[n] = P(sigma_{n+1}), [nt, nrec].
    With a store, u_n and w_n of every n = k, 2k, ... (n <= nt) go to slot n / k - 1."""
    nrec = shot.rec_index.numel()
    rec = torch.zeros((nt, nrec), dtype=torch.float32, device=P.device)
    P.reset()
    P.prime()
    for j in range(nt):
        snap = store.desc((j + 1) // k - 1) if store is not None and (j + 1) % k == 0 else None
        P.step_fwd_fused(shot.src_index, P.dt * float(shot.wavelet[j]), snap)
        s1, s2 = P.sigma
        call("tti_gather_pressure", s1.data_ptr(), s2.data_ptr(), shot.rec_index.data_ptr(), nrec,
             rec[j].data_ptr(), _stream())
    return rec


def gradient(P: TTIPropagator, shot: Shot, d_obs: torch.Tensor, nt: int, k: int = 1, snap_mode: int = 2):
    """Misfit J and its gradient dJ/d epsilon (FP64, interior shape [nz, ny, nx]) for the shot; snapshots every k
    steps in `snap_mode` (TTI_SNAP_FP32 = 1, FP16 = 2, BF16 = 3). Returns (J, grad, d_syn)."""
    if P.sponge is not None:
        raise ValueError("the gradient pipeline is exact for the unit sponge only")
    g = P.grid
    store = SnapshotStore(g, snap_mode, nt // k, P.device)
    d_syn = forward(P, shot, nt, store, k)
    res = (d_syn - d_obs).contiguous()
    J = 0.5 * float(torch.sum(res.double() ** 2))
    grad = torch.zeros((g.nz, g.ny, g.nx), dtype=torch.float64, device=P.device)
    P.reset()
    nrec = shot.rec_index.numel()
    for i in range(nt):
        n = nt - i  # this adjoint step yields mu_n
        snap = store.desc(n // k - 1) if n % k == 0 else None
        P.step_adj_fused(snap, grad if snap is not None else None)
        s1o, s2o = P.sigma  # the sigma the step wrote (adjoint sigma of the next step)
        call("tti_inject_residual", P.u.data_ptr(), P.w.data_ptr(), s1o.data_ptr(), s2o.data_ptr(),
             shot.rec_index.data_ptr(), shot.rec_ijk.data_ptr(), nrec, res[n - 1].data_ptr(), P.dt, None, None, None,
             *g.ranges, snap, grad.data_ptr() if snap is not None else None, _stream())
    grad *= P.dt * k
    return J, grad, d_syn
