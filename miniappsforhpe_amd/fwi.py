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


class RhoImage:
    """Density image of the FWI gradient: at every snapshot step, after the adjoint step and the residual
    injection, tti_rho_image adds -b sum[S D+ Dfwd + (D- S) Dfwd] (S of the adjoint sigma, Dfwd the forward derivatives
    of the stored sigma) to an FP64 image; -dt k times the sum is dJ/d rho. Holds the 8 zeroed work volumes."""

    def __init__(self, P: TTIPropagator):
        g = P.grid
        self.grid = g
        self.work = torch.zeros((8,) + g.shape, dtype=torch.float32, device=P.device)
        self.grad = torch.zeros((g.nz, g.ny, g.nx), dtype=torch.float64, device=P.device)

    def add(self, P: TTIPropagator, snap):
        g = P.grid
        s1, s2 = P.sigma
        call("tti_rho_image", s1.data_ptr(), s2.data_ptr(), *P._mptrs(), *g.ranges, g.ldx, g.ldy, P.dx, P.dy, P.dz,
             snap, self.work.data_ptr(), self.grad.data_ptr(), _stream())


def _adj_steps(P: TTIPropagator, shot: Shot, res: torch.Tensor, grad: torch.Tensor, hi: int, lo: int,
               store: SnapshotStore, k: int, slot0: int = 0, rho: RhoImage | None = None):
    """Adjoint steps yielding mu_n for n = hi down to lo + 1, each followed by the residual injection of d_{n-1};
    the imaging condition reads store slot n / k - 1 - slot0 at n divisible by k (and, with `rho`, the density image
    there as well, from the adjoint sigma after the injection)."""
    g = P.grid
    nrec = shot.rec_index.numel()
    for n in range(hi, lo, -1):
        snap = store.desc(n // k - 1 - slot0) if n % k == 0 else None
        P.step_adj_fused(snap, grad if snap is not None else None)
        s1o, s2o = P.sigma  # the sigma the step wrote (adjoint sigma of the next step)
        call("tti_inject_residual", P.u.data_ptr(), P.w.data_ptr(), s1o.data_ptr(), s2o.data_ptr(),
             shot.rec_index.data_ptr(), shot.rec_ijk.data_ptr(), nrec, res[n - 1].data_ptr(), P.dt, None, None, None,
             *g.ranges, snap, grad.data_ptr() if snap is not None else None, _stream())
        if rho is not None and snap is not None:
            rho.add(P, snap)


def gradient(P: TTIPropagator, shot: Shot, d_obs: torch.Tensor, nt: int, k: int = 1, snap_mode: int = 2,
             rho: bool = False):
    """Misfit J and its gradient dJ/d epsilon (FP64, interior shape [nz, ny, nx]) for the shot; snapshots every k
    steps in `snap_mode` (TTI_SNAP_FP32 = 1, FP16 = 2, BF16 = 3). Returns (J, grad, d_syn); with rho=True the
    snapshots also hold sigma and the density image dJ/d rho (FP64) is returned fourth."""
    if P.sponge is not None:
        raise ValueError("the gradient pipeline is exact for the unit sponge only")
    g = P.grid
    store = SnapshotStore(g, snap_mode, nt // k, P.device, sigma=rho)
    d_syn = forward(P, shot, nt, store, k)
    res = (d_syn - d_obs).contiguous()
    J = 0.5 * float(torch.sum(res.double() ** 2))
    grad = torch.zeros((g.nz, g.ny, g.nx), dtype=torch.float64, device=P.device)
    img = RhoImage(P) if rho else None
    P.reset()
    _adj_steps(P, shot, res, grad, nt, 0, store, k, rho=img)
    grad *= P.dt * k
    if rho:
        img.grad *= -P.dt * k
        return J, grad, d_syn, img.grad
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
    """A recursive checkpoint schedule. `lengths` are the segment lengths of each level, outermost first (multiples of
    k, decreasing); () stores every snapshot. A segment of level i is cut into segments of level i+1; the innermost
    segments hold their snapshots. Counts are exact for this (nt, k)."""
    nt: int
    k: int
    lengths: tuple
    checkpoints: int       # checkpoint buffers (sigma_H, sigma_V, u, w) held at once, all levels
    snapshots: int         # snapshot slots (the longest innermost segment's)
    nbytes: int            # checkpoints + snapshot slots
    recomputed_steps: int  # forward steps run again (beyond the nt of the first forward)


def _held(m: int, innermost: bool) -> int:
    """Checkpoints a segment with m children saves: every child after the first (whose start is the segment's own),
    except the last when the children are innermost (the forward goes on into it and its adjoint follows at once);
    a last child with children of its own needs its start again for its first child."""
    return max(m - 2, 0) if innermost else max(m - 1, 0)


def _walk(nt: int, lengths) -> tuple[int, int]:
    """Dry run of the schedule of `gradient_checkpointed`: (checkpoint buffers held at once, recomputed forward
    steps). Every child but the last is re-run from its checkpoint when its turn comes."""
    rec = 0

    def seg(level, a, b, first):
        nonlocal rec
        if level == len(lengths):
            rec += 0 if first else b - a
            return 0
        L = lengths[level]
        kids = [(s, min(s + L, b)) for s in range(a, b, L)]
        rec += 0 if first else kids[-1][0] - a
        deeper = 0
        for i in range(len(kids) - 1, -1, -1):
            deeper = max(deeper, seg(level + 1, *kids[i], first and i == len(kids) - 1))
        return _held(len(kids), level + 1 == len(lengths)) + deeper

    return seg(0, 0, nt, True), rec


def plan_checkpoints(grid, nt: int, k: int, mode: int, budget_bytes: int, max_levels: int = 2) -> CheckpointPlan:
    """Checkpoint schedule with the fewest recomputed forward steps among those whose snapshot slots plus checkpoints
    fit `budget_bytes` of device memory: every snapshot stored if that fits, else the best one- or two-level
    schedule (segment lengths multiples of k). Raises ValueError when none fits."""
    return plan_checkpoints_bytes(snapshot_bytes(grid, mode), checkpoint_bytes(grid), nt, k, budget_bytes,
                                  max_levels)


def plan_checkpoints_bytes(snap: int, ck: int, nt: int, k: int, budget_bytes: int,
                           max_levels: int = 2) -> CheckpointPlan:
    """plan_checkpoints for given sizes of one snapshot slot and one checkpoint (bytes): a batch of shots that step
    together shares one schedule, with the sums of its shots' sizes."""
    if nt < 1 or k < 1:
        raise ValueError("nt and k must be positive")
    ceil = lambda a, b: -(-a // b)
    top = k * ceil(nt, k)
    if (top // k) * snap <= budget_bytes:
        held, rec = _walk(nt, ())
        return CheckpointPlan(nt, k, (), held, top // k, (top // k) * snap, rec)
    best = None  # (recomputed steps, bytes, lengths)
    for L1 in range(k, top, k):
        m1 = ceil(nt, L1)
        mem = _held(m1, True) * ck + (L1 // k) * snap
        if mem <= budget_bytes:
            # the first forward reaches the last segment; every earlier segment is run once more
            cand = ((m1 - 1) * L1, mem, (L1,))
            best = cand if best is None or cand < best else best
    if max_levels >= 2:
        for L1 in range(2 * k, top, k):
            m1 = ceil(nt, L1)
            base = _held(m1, False) * ck
            if base + snap > budget_bytes:
                continue
            last1 = nt - (m1 - 1) * L1
            for L2 in range(k, L1, k):
                mem = base + _held(ceil(L1, L2), True) * ck + (L2 // k) * snap
                if mem > budget_bytes:
                    continue
                # an earlier top segment re-runs its forward to its last child's start, then every child; in the
                # last top segment (reached by the first forward) only the children before its last are re-run
                rec = (m1 - 1) * ((ceil(L1, L2) - 1) * L2 + L1) + (ceil(last1, L2) - 1) * L2
                cand = (rec, mem, (L1, L2))
                best = cand if best is None or cand < best else best
    if best is None:
        raise ValueError(f"snapshot budget {budget_bytes} B: no checkpoint schedule with up to {max_levels} levels "
                         f"fits (one snapshot {snap} B, one checkpoint {ck} B)")
    lengths = best[2]
    held, rec = _walk(nt, lengths)
    slots = lengths[-1] // k
    return CheckpointPlan(nt, k, lengths, held, slots, held * ck + slots * snap, rec)


def check_plan(plan: CheckpointPlan, nt: int, k: int) -> tuple:
    """Validates a plan's segment lengths for (nt, k); returns them as a tuple."""
    lengths = tuple(plan.lengths)
    if any(L % k or L < k for L in lengths) or any(b >= a for a, b in zip(lengths, lengths[1:])):
        raise ValueError("segment lengths must be decreasing positive multiples of k")
    return lengths


def run_schedule(nt: int, lengths: tuple, fwd, adj, first_done, save, restore, release) -> None:
    """Executes a recursive checkpoint schedule (see gradient_checkpointed) through callbacks, so that one shot and a
    batch of shots share it:
      fwd(a, b, first, stored, save_cb)  forward steps a..b-1 from the forward state; first: the first forward over
                                         these steps (record the receivers); stored: innermost segment (store the
                                         snapshots of [a, b) with slot base a // k); save_cb(j) runs before step j
      adj(hi, lo)                        adjoint steps yielding mu_n for n = hi .. lo + 1, imaging from the store
      first_done()                       the first forward reached nt: residuals and misfits
      save() -> id                       copies the forward state into a free checkpoint buffer
      restore(id | None)                 forward state from a checkpoint (None: rest)
      release(id)                        frees a checkpoint buffer
    The forward state starts at rest (the caller resets and primes it)."""

    def seg(level, a, b, first, cont, start):
        if not cont:
            restore(start)
        if level == len(lengths):
            fwd(a, b, first, True, None)
            if b == nt and first:
                first_done()
            adj(b, a)
            return
        L = lengths[level]
        kids = [(s, min(s + L, b)) for s in range(a, b, L)]
        last = len(kids) - 1
        keep_last = level + 1 < len(lengths)  # the last child's start is needed again by its own first child
        mine = {}

        def save_cb(j):
            i = (j - a) // L
            if (j - a) % L == 0 and 0 < i and (i < last or keep_last):
                mine[i] = save()

        fwd(a, kids[-1][0], first, False, save_cb)
        if keep_last and last > 0:
            save_cb(kids[-1][0])  # the forward above stops before the last child's start
        for i in range(last, -1, -1):
            if i == last:
                seg(level + 1, *kids[i], first, True, mine.get(i, start if i == 0 else None))
            else:
                seg(level + 1, *kids[i], False, False, start if i == 0 else mine[i])
            if i in mine:
                release(mine.pop(i))

    try:
        seg(0, 0, nt, True, True, None)
    finally:
        # seg is recursive, so its closure refers to itself: without this the cycle would keep every callback, and
        # the device buffers they hold (snapshot store, checkpoints, the re-run propagator), alive until Python's
        # cyclic collector happens to run
        seg = None


def gradient_checkpointed(P: TTIPropagator, shot: Shot, d_obs: torch.Tensor, nt: int, k: int = 1,
                          snap_mode: int = 2, budget_bytes: int | None = None, plan: CheckpointPlan | None = None,
                          rho: bool = False):
    """`gradient` under a device-memory budget for snapshots (recursive checkpointing with recomputation).

    The first forward runs once from rest over all nt steps, recording the receivers. At every level a segment's
    forward saves a checkpoint (sigma_H, sigma_V, u, w) at the start of each child segment except the first (its
    start is the segment's own) and the last (the forward goes on into it live); the innermost segments store their
    snapshots. The adjoint then takes the children last to first: the last right after the live forward, each
    earlier one after restoring its checkpoint into a second set of state buffers (six volumes, outside the budget)
    and re-running its forward. Snapshots and imaging are bit-identical to `gradient` (the fused forward is
    deterministic), so the result is too; plan.recomputed_steps is the extra cost. Returns (J, grad, d_syn, plan), and
    with rho=True the density image after the plan (the plan's snapshot sizes do not count the sigma fields)."""
    if P.sponge is not None:
        raise ValueError("the gradient pipeline is exact for the unit sponge only")
    g = P.grid
    if plan is None:
        if budget_bytes is None:
            raise ValueError("give budget_bytes or plan")
        plan = plan_checkpoints(g, nt, k, snap_mode, budget_bytes)
    lengths = check_plan(plan, nt, k)
    held, _ = _walk(nt, lengths)
    inner = lengths[-1] if lengths else k * (-(-nt // k))
    store = SnapshotStore(g, snap_mode, max(inner // k, 1), P.device, sigma=rho)
    img = RhoImage(P) if rho else None
    pool = [[torch.empty(g.shape, dtype=torch.float32, device=P.device) for _ in range(4)] for _ in range(held)]
    free = list(range(held))

    # Q runs every forward (the first one included); P runs the adjoint
    Q = TTIPropagator(g, P.models, P.dx, P.dy, P.dz, P.dt, device=P.device, zblock=P.zblock, strip=P.strip)

    def state(X):
        s1, s2 = X.sigma
        return [s1, s2, X.u, X.w]

    def save():
        i = free.pop()
        for dst, src in zip(pool[i], state(Q)):
            dst.copy_(src)
        return i

    def restore(ck):
        Q.reset()
        if ck is None:
            Q.prime()
        else:
            for dst, src in zip(state(Q), pool[ck]):
                dst.copy_(src)
            Q.primed = True

    rec = torch.zeros((nt, shot.rec_index.numel()), dtype=torch.float32, device=P.device)
    out = {}
    grad = torch.zeros((g.nz, g.ny, g.nx), dtype=torch.float64, device=P.device)

    def fwd(a, b, first, stored, save_cb):
        _fwd_steps(Q, shot, a, b, rec if first else None, store if stored else None, k, slot0=a // k,
                   checkpoint=save_cb)

    def first_done():
        res = (rec - d_obs).contiguous()
        out["res"], out["J"] = res, 0.5 * float(torch.sum(res.double() ** 2))
        P.reset()

    def adj(hi, lo):
        _adj_steps(P, shot, out["res"], grad, hi, lo, store, k, slot0=lo // k, rho=img)

    Q.reset()
    Q.prime()
    run_schedule(nt, lengths, fwd, adj, first_done, save, restore, free.append)
    grad *= P.dt * k
    if rho:
        img.grad *= -P.dt * k
        return out["J"], grad, rec, plan, img.grad
    return out["J"], grad, rec, plan
