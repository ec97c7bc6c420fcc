"""Many shots, uneven apertures, one launch per time step: the batched FWI gradient on the fused gfx950 steps.

Each shot owns a subvolume ("aperture") cut from one large model, with its own padded volumes, snapshot store,
receivers and FP64 image. `BatchFWI` builds the shot table of include/tti_batch.h once; a forward pass and an
adjoint pass are then one `tti_batch_forward` / `tti_batch_adjoint` call each, replayed from a captured hipGraph after
the first call. Every shot's records, misfit and image are bit-identical to running that shot alone through
`fwi.gradient` (the same tile bodies run on the same data); the per-shot images are summed into the model-sized
gradient in FP64 in shot order.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import fwi
from ._lib import TTIShot, call, lib
from .grid import Grid
from .propagator import SNAP_NONE, SnapshotStore


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


@dataclass
class Aperture:
    """A shot on the model's interior grid: the subvolume [x0, x0+nx) x [y0, y0+ny) x [z0, z0+nz), the source point
    and the receiver points (model interior coordinates x, y, z; receivers distinct and inside the subvolume)."""
    x0: int
    y0: int
    z0: int
    nx: int
    ny: int
    nz: int
    src: tuple[int, int, int]
    receivers: np.ndarray  # (nrec, 3) int

    def grid(self) -> Grid:
        return Grid(self.nx, self.ny, self.nz)


def seeded_apertures(model_grid: Grid, nshots: int, seed: int, lo: int = 128, hi: int = 320,
                     nrec: int = 64) -> list[Aperture]:
    """Uneven apertures drawn from a seeded generator: per shot the sizes nx, ny, nz uniform in [lo, hi] (clipped to
    the model), the origin uniform over the positions that keep the aperture inside the model, the source at the
    centre of the aperture in x-y, 8 planes below its top, and `nrec` receivers on an x line through the centre, 4
    planes below the top. The 64-shot layout of seed 1004 from a 1024 x 1024 x 400 model is the batched workload."""
    rng = np.random.default_rng(seed)
    aps = []
    for _ in range(nshots):
        nx, ny, nz = (min(int(v), m) for v, m in zip(rng.integers(lo, hi + 1, 3),
                                                      (model_grid.nx, model_grid.ny, model_grid.nz)))
        x0, y0, z0 = (int(rng.integers(0, m - n + 1)) for m, n in ((model_grid.nx, nx), (model_grid.ny, ny),
                                                                 (model_grid.nz, nz)))
        step = max(1, (nx - 8) // nrec)
        rec = np.array([(x0 + 4 + j * step, y0 + ny // 2, z0 + 4) for j in range(min(nrec, nx - 8))])
        aps.append(Aperture(x0, y0, z0, nx, ny, nz, (x0 + nx // 2, y0 + ny // 2, z0 + 8), rec))
    return aps


def cut_models(models, model_grid: Grid, ap: Aperture) -> list[torch.Tensor]:
    """The 8 model volumes of an aperture: the window with its 4-point ghost rim taken from the padded model volumes
    (real neighbour values inside the model, the model's replicated ghosts at its faces), x padding replicated."""
    g = model_grid.ghost
    sg = ap.grid()
    if not (0 <= ap.x0 and ap.x0 + ap.nx <= model_grid.nx and 0 <= ap.y0 and ap.y0 + ap.ny <= model_grid.ny
            and 0 <= ap.z0 and ap.z0 + ap.nz <= model_grid.nz):
        raise ValueError(f"aperture {ap} outside the model grid {model_grid}")
    out = []
    for m in models:
        w = m[ap.z0 + g - sg.ghost: ap.z0 + g + ap.nz + sg.ghost,
              ap.y0 + g - sg.ghost: ap.y0 + g + ap.ny + sg.ghost,
              ap.x0 + g - sg.ghost: ap.x0 + g + ap.nx + sg.ghost]
        pad = sg.ldx - w.shape[2]
        w = torch.nn.functional.pad(w[None, None], (0, pad, 0, 0, 0, 0), mode="replicate")[0, 0]
        out.append(w.contiguous())
    return out


class _ShotBuffers:
    def __init__(self, ap: Aperture, models, nt: int, k: int, snap_mode: int, wavelet_amp: torch.Tensor, device,
                 slots: int | None = None):
        self.ap, self.grid = ap, ap.grid()
        g = self.grid
        z = lambda: torch.zeros(g.shape, dtype=torch.float32, device=device)
        self._z = z
        self.s1, self.s2 = [z(), z()], [z(), z()]
        self.u, self.w = z(), z()
        self.q = None  # second state set (s1 pair, s2 pair, u, w) for the forward re-runs of a checkpointed gradient
        self.models = models
        gg = g.ghost
        sx, sy, sz = (c - o + gg for c, o in zip(ap.src, (ap.x0, ap.y0, ap.z0)))
        self.src_index = int(np.ravel_multi_index((sz, sy, sx), g.shape))
        self.amp = wavelet_amp
        rec = np.asarray(ap.receivers, dtype=np.int64).reshape(-1, 3)
        ijk = rec - np.array([ap.x0, ap.y0, ap.z0]) + gg
        if len(ijk) and (ijk.min() < gg or (ijk - gg >= np.array([g.nx, g.ny, g.nz])).any()):
            raise ValueError("receivers must lie inside the aperture")
        idx = np.array([np.ravel_multi_index((c[2], c[1], c[0]), g.shape) for c in ijk], dtype=np.int64)
        if len(np.unique(idx)) != len(idx):
            raise ValueError("receivers must be distinct points")
        self.nrec = len(idx)
        self.rec_index = torch.from_numpy(idx).to(device)
        self.rec_ijk = torch.from_numpy(ijk.astype(np.int32).ravel()).to(device)
        self.records = torch.zeros((nt, max(self.nrec, 1)), dtype=torch.float32, device=device)
        self.residual = torch.zeros_like(self.records)
        cap = max(nt // k, 1) if slots is None else slots
        self.store = SnapshotStore(g, snap_mode, cap, device) if snap_mode != SNAP_NONE else None
        self.grad = torch.zeros((g.nz, g.ny, g.nx), dtype=torch.float64, device=device)

    def state(self, which: int = 0):
        """State volumes (s1 pair, s2 pair, u, w) of set 0 (adjoint, and every pass of an uncheckpointed gradient) or
        set 1 (forward re-runs of a checkpointed gradient, allocated on first use)."""
        if which == 0:
            return self.s1, self.s2, self.u, self.w
        if self.q is None:
            z = self._z
            self.q = ([z(), z()], [z(), z()], z(), z())
        return self.q

    def desc(self, nt: int, zblock: int, strip: int = 0, which: int = 0) -> TTIShot:
        s1, s2, u, w = self.state(which)
        d = TTIShot()
        d.field1[0], d.field1[1] = s1[0].data_ptr(), s1[1].data_ptr()
        d.field2[0], d.field2[1] = s2[0].data_ptr(), s2[1].data_ptr()
        d.field3, d.field4 = u.data_ptr(), w.data_ptr()
        for i, m in enumerate(self.models):
            d.model[i] = m.data_ptr()
        for i in range(3):
            d.AB[i] = None
        (d.ixbeg, d.ixend, d.iybeg, d.iyend, d.izbeg, d.izend) = self.grid.ranges
        d.ldimx, d.ldimy = self.grid.ldx, self.grid.ldy
        d.src_index = self.src_index
        d.src_amplitude = self.amp.data_ptr()
        d.nt_max = nt
        if self.store is not None:
            d.snap = self.store._desc[0]
            d.snap_slot_stride = self.store.u[0].numel()
            d.snap_slots = self.store.capacity
        else:
            d.snap.mode = SNAP_NONE
        d.grad = self.grad.data_ptr()
        d.rec_index = self.rec_index.data_ptr()
        d.rec_ijk = self.rec_ijk.data_ptr()
        d.nrec = self.nrec
        d.records = self.records.data_ptr()
        d.residual = self.residual.data_ptr()
        d.zblock = zblock
        d.strip = strip
        return d


class BatchFWI:
    """FWI gradient of many shots through the batched fused steps.

    models: the 8 padded model volumes of `model_grid`; apertures: one per shot; wavelet: host float32 [nt] shared by
    the shots (dt * wavelet[n] is injected at step n, as fwi.forward does); snapshots every k steps in `snap_mode`.
    """

    def __init__(self, models, model_grid: Grid, apertures: list[Aperture], wavelet: torch.Tensor, dx: float,
                 dy: float, dz: float, dt: float, nt: int, k: int = 1, snap_mode: int = 2, zblock: int = 128,
                 strip: int = 0, device="cuda", slots: int | None = None):
        self.device = torch.device(device)
        self.model_grid, self.nt, self.k, self.dt = model_grid, nt, k, float(dt)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.snap_mode, self.zblock = snap_mode, zblock
        # source amplitudes rounded as the single-shot path rounds dt * wavelet[n] (FP64 product, then FP32)
        amp = (np.float64(self.dt) * wavelet.double().numpy()).astype(np.float32)
        self.amp = torch.from_numpy(amp).to(self.device)
        self.shots = [_ShotBuffers(ap, cut_models(models, model_grid, ap), nt, k, snap_mode, self.amp, self.device,
                                   slots) for ap in apertures]
        self.slots = max(nt // k, 1) if slots is None else slots
        self._h = None
        self._hq = None
        self.retile(zblock, strip)

    def retile(self, zblock: int, strip: int = 0):
        """Rebuilds the shot table with `zblock` z planes per thread block and the block schedule `strip` (x tiles per
        strip, 0 = full rows); captured graphs are dropped."""
        self.close()
        self.zblock, self.strip = zblock, strip
        self._descs = (TTIShot * len(self.shots))(*[s.desc(self.nt, zblock, strip) for s in self.shots])
        h = C.c_void_p()
        call("tti_batch_create", self._descs, len(self.shots), self.dt, self.dx, self.dy, self.dz, C.byref(h))
        self._h = h

    def _forward_handle(self):
        """The batch over the second state set (forward re-runs of a checkpointed gradient), built on first use."""
        if self._hq is None:
            self._qdescs = (TTIShot * len(self.shots))(*[s.desc(self.nt, self.zblock, self.strip, which=1)
                                                          for s in self.shots])
            h = C.c_void_p()
            call("tti_batch_create", self._qdescs, len(self.shots), self.dt, self.dx, self.dy, self.dz, C.byref(h))
            self._hq = h
        return self._hq

    def step_fwd(self, parity: int, it: int, slot: int = -1):
        """One batched forward step (all shots) + receivers; see tti_batch_fwd_step."""
        call("tti_batch_fwd_step", self._h, parity, it, slot, _stream())

    def step_adj(self, parity: int, n: int, slot: int = -1):
        """One batched adjoint step (all shots) + residual injection; see tti_batch_adj_step."""
        call("tti_batch_adj_step", self._h, parity, n, slot, _stream())

    def close(self):
        for name in ("_h", "_hq"):
            h = getattr(self, name, None)
            if h is not None and h.value is not None:
                torch.cuda.synchronize(self.device)
                call("tti_batch_destroy", h)
            setattr(self, name, None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tiles(self) -> int:
        return lib().tti_batch_tiles(self._h)

    def interior_points(self) -> int:
        return sum(s.grid.npoints for s in self.shots)

    def forward(self, use_graph: bool = True):
        """Forward pass of every shot from rest; records land in each shot's `records` [nt, nrec]."""
        call("tti_batch_forward", self._h, self.nt, self.k, int(use_graph), _stream())

    def adjoint(self, use_graph: bool = True):
        """Adjoint pass of every shot from rest, injecting each shot's `residual` and accumulating its image."""
        call("tti_batch_adjoint", self._h, self.nt, self.k, int(use_graph), _stream())

    def forward_records(self) -> list[torch.Tensor]:
        """A forward pass of every shot from rest without snapshots (one launch per step, no graph); returns a copy of
        each shot's records [nt, nrec]. Observed data of synthetic tests are made this way."""
        st = _stream()
        call("tti_batch_reset", self._h, st)
        for j in range(self.nt):
            call("tti_batch_fwd_step", self._h, j & 1, j, -1, st)
        return [s.records[:, : s.nrec].clone() for s in self.shots]

    def snapshot_slot_bytes(self) -> int:
        """Device bytes of one snapshot slot of every shot (u and w)."""
        return sum(s.store.nbytes() // s.store.capacity for s in self.shots)

    def checkpoint_bytes(self) -> int:
        """Device bytes of one checkpoint of every shot (sigma_H, sigma_V, u, w)."""
        return sum(4 * s.grid.nbytes() for s in self.shots)

    def plan(self, budget_bytes: int) -> fwi.CheckpointPlan:
        """The checkpoint schedule (shared by all shots, which step together) with the fewest re-run steps whose
        checkpoints and snapshot slots fit `budget_bytes`."""
        return fwi.plan_checkpoints_bytes(self.snapshot_slot_bytes(), self.checkpoint_bytes(), self.nt, self.k,
                                          budget_bytes)

    def _images(self):
        mg = self.model_grid
        total = torch.zeros((mg.nz, mg.ny, mg.nx), dtype=torch.float64, device=self.device)
        grads = []
        for s in self.shots:
            s.grad *= self.dt * self.k
            ap = s.ap
            total[ap.z0: ap.z0 + ap.nz, ap.y0: ap.y0 + ap.ny, ap.x0: ap.x0 + ap.nx] += s.grad
            grads.append(s.grad)
        return grads, total

    def _residuals(self, d_obs):
        J = []
        for s, d in zip(self.shots, d_obs):
            r = s.records[:, : s.nrec] - d
            s.residual[:, : s.nrec].copy_(r)
            J.append(0.5 * float(torch.sum(r.double() ** 2)))
            s.grad.zero_()
        return J

    def gradient_checkpointed(self, d_obs: list[torch.Tensor], plan: fwi.CheckpointPlan, progress=None):
        """`gradient` with the shots' snapshots under a checkpoint schedule (recursive checkpointing with
        recomputation, the executor of fwi.run_schedule). All shots step together, so one schedule serves the batch:
        a checkpoint holds every shot's sigma_H, sigma_V, u, w; the forward steps run on a second state set, the
        adjoint on the first. Each shot's store needs plan.snapshots slots (the `slots` of the constructor). Records,
        misfits and images are bit-identical to `gradient` (the fused steps are deterministic). `progress(lo, hi)`,
        when given, runs after each innermost segment's adjoint steps hi .. lo + 1 have finished on the device (the
        stream is synchronised for it)."""
        k, nt = self.k, self.nt
        lengths = fwi.check_plan(plan, nt, k)
        if plan.nt != nt or plan.k != k:
            raise ValueError("the plan is for another (nt, k)")
        inner = lengths[-1] // k if lengths else -(-nt // k)
        if inner > self.slots:
            raise ValueError(f"the plan needs {inner} snapshot slots per shot; the batch has {self.slots}")
        hq, hp = self._forward_handle(), self._h
        st = _stream()
        pool, free = [], []
        cur = {"j": 0, "J": None}

        def save():
            if not free:
                pool.append([[torch.empty_like(s.u) for _ in range(4)] for s in self.shots])
                free.append(len(pool) - 1)
            i = free.pop()
            par = cur["j"] & 1
            for buf, s in zip(pool[i], self.shots):
                s1, s2, u, w = s.state(1)
                for dst, src in zip(buf, (s1[par], s2[par], u, w)):
                    dst.copy_(src)
            pool[i].append(par)
            return i

        def restore(ck):
            call("tti_batch_reset", hq, st)
            if ck is None:
                return
            par = pool[ck][-1]
            for buf, s in zip(pool[ck][:-1], self.shots):
                s1, s2, u, w = s.state(1)
                for dst, src in zip((s1[par], s2[par], u, w), buf):
                    dst.copy_(src)

        def release(i):
            pool[i].pop()
            free.append(i)

        def fwd(a, b, first, stored, save_cb):
            for j in range(a, b):
                cur["j"] = j
                if save_cb is not None:
                    save_cb(j)
                n = j + 1
                slot = n // k - 1 - a // k if stored and n % k == 0 else -1
                call("tti_batch_fwd_step", hq, j & 1, j, slot, st)
            cur["j"] = b  # the forward state is now the one before step b

        def first_done():
            cur["J"] = self._residuals(d_obs)
            call("tti_batch_reset", hp, st)

        def adj(hi, lo):
            for n in range(hi, lo, -1):
                slot = n // k - 1 - lo // k if n % k == 0 else -1
                call("tti_batch_adj_step", hp, (nt - n) & 1, n, slot, st)
            if progress is not None:
                torch.cuda.synchronize()
                progress(lo, hi)

        call("tti_batch_reset", hq, st)
        fwi.run_schedule(nt, lengths, fwd, adj, first_done, save, restore, release)
        grads, total = self._images()
        return cur["J"], grads, total

    def gradient(self, d_obs: list[torch.Tensor], use_graph: bool = True):
        """Misfits J_s, per-shot gradients dJ_s/d epsilon (FP64, aperture interior) and their FP64 sum on the model
        interior grid (shot order). d_obs[s] is [nt, nrec_s] on the device."""
        self.forward(use_graph)
        J = self._residuals(d_obs)
        self.adjoint(use_graph)
        grads, total = self._images()
        return J, grads, total
