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

import math
from dataclasses import dataclass

import torch

from ._lib import call
from .grid import roundup
from .propagator import SNAP_FP16, SnapshotStore, TTIPropagator


@dataclass
class Shot:
    src_index: int            # linear index of the source in the padded volume
    wavelet: torch.Tensor     # [nt] host float32; dt * wavelet[n] is added to u and w at step n
    rec_index: torch.Tensor   # [nrec] int64 linear indices (device)
    rec_ijk: torch.Tensor     # [3 * nrec] int32 array coordinates x, y, z of each receiver (device, distinct points)


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _fwd_steps(P: TTIPropagator, shot: Shot, a: int, b: int, rec: torch.Tensor | None,
               store: SnapshotStore | None, k: int, slot0: int = 0, checkpoint=None):
    """Fused forward steps j = a..b-1 from P's current state. u_n, w_n of n = j + 1 divisible by k go to store slot
    n / k - 1 - slot0; rec[j] receives d_j; checkpoint(j) is called before step j."""
    nrec = shot.rec_index.numel()
    for j in range(a, b):
        if checkpoint is not None:
            checkpoint(j)
        n = j + 1
        snap = store.desc(n // k - 1 - slot0) if store is not None and n % k == 0 else None
        P.step_fwd_fused(shot.src_index, P.dt * float(shot.wavelet[j]), snap)
        if rec is not None:
            s1, s2 = P.sigma
            call("tti_gather_pressure", s1.data_ptr(), s2.data_ptr(), shot.rec_index.data_ptr(), nrec,
                 rec[j].data_ptr(), _stream())


def forward(P: TTIPropagator, shot: Shot, nt: int, store: SnapshotStore | None = None, k: int = 1) -> torch.Tensor:
    """Runs nt fused forward steps from rest and returns the receiver pressures d[n] = P(sigma_{n+1}), [nt, nrec].
    With a store, u_n and w_n of every n = k, 2k, ... (n <= nt) go to slot n / k - 1."""
    rec = torch.zeros((nt, shot.rec_index.numel()), dtype=torch.float32, device=P.device)
    P.reset()
    P.prime()
    _fwd_steps(P, shot, 0, nt, rec, store, k)
    return rec


def _adj_steps(P: TTIPropagator, shot: Shot, res: torch.Tensor, grad: torch.Tensor, hi: int, lo: int,
               store: SnapshotStore, k: int, slot0: int = 0):
    """Adjoint steps yielding mu_n for n = hi down to lo + 1, each followed by the residual injection of d_{n-1};
    the imaging condition reads store slot n / k - 1 - slot0 at n divisible by k."""
    g = P.grid
    nrec = shot.rec_index.numel()
    for n in range(hi, lo, -1):
        snap = store.desc(n // k - 1 - slot0) if n % k == 0 else None
        P.step_adj_fused(snap, grad if snap is not None else None)
        s1o, s2o = P.sigma  # the sigma the step wrote (adjoint sigma of the next step)
        call("tti_inject_residual", P.u.data_ptr(), P.w.data_ptr(), s1o.data_ptr(), s2o.data_ptr(),
             shot.rec_index.data_ptr(), shot.rec_ijk.data_ptr(), nrec, res[n - 1].data_ptr(), P.dt, None, None, None,
             *g.ranges, snap, grad.data_ptr() if snap is not None else None, _stream())


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
    _adj_steps(P, shot, res, grad, nt, 0, store, k)
    grad *= P.dt * k
    return J, grad, d_syn


# ---- snapshot budget: checkpoints + recomputation --------------------------------------------------------------------

def snapshot_bytes(grid, mode: int) -> int:
    """Device bytes of one (u, w) snapshot in `mode` (FP32 4 B, FP16 2 B + one FP32 scale per 64 points, BF16 2 B)."""
    ldx = roundup(grid.nx, 64)
    pts = grid.nz * grid.ny * ldx
    per = pts * (4 if mode == 1 else 2)
    if mode == SNAP_FP16:
        per += (pts // 64) * 4
    return 2 * per


def checkpoint_bytes(grid) -> int:
    """Device bytes of one forward checkpoint: sigma_H, sigma_V, u, w as full padded FP32 volumes."""
    return 4 * grid.nbytes()


@dataclass
class CheckpointPlan:
    segment: int          # forward steps per segment (a multiple of k)
    segments: int         # ceil(nt / segment)
    checkpoints: int      # stored forward states: the starts of segments 1 .. segments-2 (segment 0 starts from
                          # rest, the last segment keeps its snapshots from the first pass)
    snapshots: int        # snapshot slots held at once (one segment's)
    nbytes: int           # checkpoints + snapshot slots
    recomputed_steps: int  # extra forward steps: every segment but the last is run twice


def plan_checkpoints(grid, nt: int, k: int, mode: int, budget_bytes: int) -> CheckpointPlan:
    """Segment length for a two-level checkpoint schedule that fits `budget_bytes` of device memory for snapshots
    plus checkpoints. The whole run in one segment (every snapshot stored, no recomputation) is taken when it fits;
    otherwise the longest segment that fits, which stores the fewest checkpoints. Raises ValueError when even the
    smallest schedule does not fit."""
    if nt < 1 or k < 1:
        raise ValueError("nt and k must be positive")
    snap, ck = snapshot_bytes(grid, mode), checkpoint_bytes(grid)
    best = None
    for seg in range(k * ((nt + k - 1) // k), 0, -k):
        segs = (nt + seg - 1) // seg
        nck = max(segs - 2, 0)
        need = nck * ck + (seg // k) * snap
        if need <= budget_bytes:
            best = CheckpointPlan(seg, segs, nck, seg // k, need, (segs - 1) * seg)
            break
    if best is None:
        low = min(max((nt + s - 1) // s - 2, 0) * ck + (s // k) * snap for s in range(k, nt + k, k))
        raise ValueError(f"snapshot budget {budget_bytes} B is below the smallest checkpoint schedule ({low} B)")
    return best


def gradient_checkpointed(P: TTIPropagator, shot: Shot, d_obs: torch.Tensor, nt: int, k: int = 1,
                          snap_mode: int = 2, budget_bytes: int | None = None, plan: CheckpointPlan | None = None):
    """`gradient` under a device-memory budget for snapshots: pass 1 runs the forward once, keeping the receiver
    records, a checkpoint (sigma_H, sigma_V, u, w) at the start of every segment but the first (which starts from
    rest) and the last, and the snapshots of the last segment; the adjoint then walks the segments backwards,
    re-running each earlier segment's forward from its checkpoint on a second set of state buffers (six more
    volumes, outside the budget) to refill the snapshot slots before its adjoint steps. Snapshots
    and imaging are bit-identical to `gradient` (the fused forward is deterministic), so the result is too; the
    cost is the recomputed steps of plan.recomputed_steps. Returns (J, grad, d_syn, plan)."""
    if P.sponge is not None:
        raise ValueError("the gradient pipeline is exact for the unit sponge only")
    g = P.grid
    if plan is None:
        if budget_bytes is None:
            raise ValueError("give budget_bytes or plan")
        plan = plan_checkpoints(g, nt, k, snap_mode, budget_bytes)
    L = plan.segment
    if L % k or L < k:
        raise ValueError("the segment length must be a positive multiple of k")
    segs = (nt + L - 1) // L
    store = SnapshotStore(g, snap_mode, L // k, P.device)
    ckpt = [[torch.empty(g.shape, dtype=torch.float32, device=P.device) for _ in range(4)]
            for _ in range(max(segs - 2, 0))]

    def state(Q):
        s1, s2 = Q.sigma
        return [s1, s2, Q.u, Q.w]

    def save(j):  # before step j: the start state of segment j / L (1 .. segs - 2)
        if j % L == 0 and j > 0:
            for dst, src in zip(ckpt[j // L - 1], state(P)):
                dst.copy_(src)

    # pass 1: records and checkpoints; the last segment keeps its snapshots
    rec = torch.zeros((nt, shot.rec_index.numel()), dtype=torch.float32, device=P.device)
    P.reset()
    P.prime()
    last = (segs - 1) * L
    _fwd_steps(P, shot, 0, last, rec, None, k, checkpoint=save)
    _fwd_steps(P, shot, last, nt, rec, store, k, slot0=last // k)
    d_syn = rec
    res = (d_syn - d_obs).contiguous()
    J = 0.5 * float(torch.sum(res.double() ** 2))

    grad = torch.zeros((g.nz, g.ny, g.nx), dtype=torch.float64, device=P.device)
    Q = None
    P.reset()
    for s in range(segs - 1, -1, -1):
        a, b = s * L, min(s * L + L, nt)
        if s < segs - 1:
            if Q is None:
                Q = TTIPropagator(g, P.models, P.dx, P.dy, P.dz, P.dt, device=P.device, zblock=P.zblock)
            Q.reset()
            if s == 0:
                Q.prime()
            else:
                for dst, src in zip(state(Q), ckpt[s - 1]):
                    dst.copy_(src)
                Q.primed = True
            _fwd_steps(Q, shot, a, b, None, store, k, slot0=a // k)
        _adj_steps(P, shot, res, grad, b, a, store, k, slot0=a // k)
    grad *= P.dt * k
    return J, grad, d_syn, plan
