"""FWI gradient: definition pinned by finite differences in the FP64 oracle, then the GPU pipeline against it.

CPU (oracle only):
* central finite differences of the misfit J along random z-profiles of the sigma-update coefficient
  B = 1 + epsilon(z) (Richardson-extrapolated) agree with <grad, delta> from the adjoint + imaging condition to 1e-7
  (FP64). A wrong residual
  sign or weight, a time shift between snapshots and adjoint states, or an adjoint that is not the transpose fails it;
* with snapshots every k = 2 steps the subsampled gradient stays within 1% of the full one (quadrature, not a bug).

GPU (fused kernels through miniappsforhpe_amd.fwi against the oracle twin on identical FP32 models):
* misfit and synthetic records to FP32 accuracy;
* FP32 snapshots, k = 1: gradient rel-L2 <= 1e-5 (measured 2.4e-7 on MI355X);
* FP16 snapshots (per-64-point power-of-two scale), k = 2: rel-L2 <= 1e-3 against the FP64 oracle with the same k
  (measured 1.0e-4);
  BF16 (unscaled, 8-bit mantissa): rel-L2 <= 3e-3 (measured 9.5e-4);
* the GPU gradient against centred finite differences of the GPU misfit along 16 random smooth directions (1e-3);
* the density image: dJ/d rho pinned by central differences in the FP64 oracle (1e-9); on the GPU against the oracle
  (FP32 snapshots 1e-5, FP16 2e-3) and against central differences of the GPU misfit along 6 density directions.
"""
import numpy as np
import pytest

import oracle as O
from fwi_oracle import OracleFWI, ricker, shot_geometry, smooth_models

N = (28, 24, 26)
NT = 90
DX = DY = DZ = 10.0
DT = 1.0e-3


def setup_oracle(seed=7):
    G = O.Geom(*N)
    m = smooth_models(G, seed)
    F = OracleFWI(G, m, DX, DY, DZ, DT)
    src, idx, ijk = shot_geometry(G)
    wav = ricker(NT, DT, 30.0) * 1e6
    # observed data: the same shot with a scaled wavelet, so the residual is not proportional to the records
    d_obs, _ = F.forward(src, 0.7 * wav, idx, NT)
    return G, m, F, src, idx, ijk, wav, d_obs


def test_gradient_matches_finite_differences():
    G, m, F, src, idx, ijk, wav, d_obs = setup_oracle()
    J0, g, _ = F.gradient(src, wav, idx, d_obs, NT, k=1)
    assert J0 > 0
    gz = g.sum(axis=(1, 2))  # dJ/d epsilon for a per-plane perturbation
    rng = np.random.default_rng(11)
    # the oracle's sponge pairs are FP32 (as the reference's float2): perturbations are multiples of 2^-13 so that
    # 1 + h delta is exact; Richardson extrapolation of two central differences removes the O(h^2) term
    for _ in range(2):
        delta = rng.integers(-8, 9, G.nz) / 8.0
        bz = lambda s: np.concatenate([np.ones(G.g), 1.0 + s * delta, np.ones(G.ldz - G.g - G.nz)])
        fd = []
        for h in (2.0**-9, 2.0**-10):
            jp = F.misfit(src, wav, idx, d_obs, NT, bz(h))
            jm = F.misfit(src, wav, idx, d_obs, NT, bz(-h))
            fd.append((jp - jm) / (2 * h))
        fd = (4 * fd[1] - fd[0]) / 3
        an = float(np.dot(gz, delta))
        assert abs(fd - an) <= 1e-7 * max(abs(fd), abs(an)), (fd, an)


def test_density_gradient_matches_finite_differences():
    """The density image of the FWI gradient (the reference's update_grad_loop_2: forward derivatives of the snapshot
    sigma, then the density-gradient operator with the adjoint sigma after the step and the residual injection,
    accumulated as -dt k sum) is dJ/d rho: Richardson-extrapolated central differences of the misfit along a random
    density direction through the whole forward time loop (rho enters only through b = 1/rho) agree to 1e-9 (FP64).
    A wrong time pairing (sigma_n instead of sigma_{n-1}), the adjoint sigma before the injection, or a wrong sign or
    scale fails it."""
    G, m, F, src, idx, ijk, wav, d_obs = setup_oracle()
    nt = 60
    J0, _, _, grho = F.gradient(src, wav[:nt], idx, d_obs[:nt], nt, k=1, rho=True)
    rng = np.random.default_rng(12)
    rho = 1.0 / m[7]
    delta = np.zeros(G.shape)
    G.interior(delta)[...] = rng.standard_normal((G.nz, G.ny, G.nx))

    def J(h):
        mm = list(m)
        mm[7] = 1.0 / (rho + h * delta)
        return OracleFWI(G, mm, DX, DY, DZ, DT).misfit(src, wav[:nt], idx, d_obs[:nt], nt)

    fd = [(J(h) - J(-h)) / (2 * h) for h in (1e-1, 5e-2)]
    fd = (4 * fd[1] - fd[0]) / 3
    an = float(np.sum(grho * G.interior(delta)))
    assert abs(fd - an) <= 1e-9 * abs(fd), (fd, an)


def test_subsampled_gradient_is_close():
    G, m, F, src, idx, ijk, wav, d_obs = setup_oracle()
    _, g1, _ = F.gradient(src, wav, idx, d_obs, NT, k=1)
    _, g2, _ = F.gradient(src, wav, idx, d_obs, NT, k=2)
    assert np.linalg.norm(g2 - g1) / np.linalg.norm(g1) < 1e-2


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
@pytest.mark.parametrize("mode,k,tol", [(1, 1, 1e-5), (2, 2, 1e-3), (3, 2, 3e-3)])
def test_gpu_gradient_vs_oracle(mode, k, tol, cuda):
    import torch

    from miniappsforhpe_amd import fwi
    from miniappsforhpe_amd.grid import Grid
    from miniappsforhpe_amd.propagator import TTIPropagator

    G, m, F, src, idx, ijk, wav, d_obs = setup_oracle()
    print(f"gradient mode={mode} k={k} grid={N} nt={NT}", flush=True)
    Jo, go, do = F.gradient(src, wav, idx, d_obs, NT, k=k)
    grid = Grid(*N)
    assert grid.shape == G.shape
    P = TTIPropagator(grid, [torch.from_numpy(a.astype(np.float32)) for a in m], DX, DY, DZ, DT, device=cuda,
                      zblock=16)
    shot = fwi.Shot(src, torch.from_numpy(wav), torch.from_numpy(idx).to(cuda), torch.from_numpy(ijk).to(cuda))
    J, g, d_syn = fwi.gradient(P, shot, torch.from_numpy(d_obs.astype(np.float32)).to(cuda), NT, k=k, snap_mode=mode)
    torch.cuda.synchronize()
    assert _rel(d_syn.cpu().numpy().astype(np.float64), do) < 2e-5
    assert abs(J - Jo) <= 1e-4 * Jo
    e = _rel(g.cpu().numpy(), go)
    print(f"gradient rel-L2 vs FP64 oracle: {e:.3e}", flush=True)
    assert e <= tol, e


@pytest.mark.gpu
def test_gpu_gradient_finite_differences_16_directions(cuda):
    """The GPU gradient (fused adjoint + FP64 imaging, FP32 snapshots, k = 1) against centred finite differences of
    the GPU misfit along 16 random smooth per-plane directions of the sigma-update coefficient: B(z) = 1 + h delta(z)
    enters through the z sponge pair (x and y pairs B = 10 so that the minimum picks the z value, A = 1 everywhere),
    delta a sum of three random cosines in multiples of 1/64 so that 1 + h delta is exact in FP32; Richardson
    extrapolation of h = 2^-7 and 2^-8 removes the O(h^2) term. Bar: 1e-3 relative per direction."""
    import torch

    from miniappsforhpe_amd import fwi
    from miniappsforhpe_amd.grid import Grid
    from miniappsforhpe_amd.propagator import TTIPropagator

    G = O.Geom(*N)
    m = [torch.from_numpy(a.astype(np.float32)) for a in smooth_models(G, 7)]
    src, idx, ijk = shot_geometry(G)
    wav = ricker(NT, DT, 30.0) * 1e6
    grid = Grid(*N)
    shot = fwi.Shot(src, torch.from_numpy(wav), torch.from_numpy(idx).to(cuda), torch.from_numpy(ijk).to(cuda))
    print(f"GPU finite differences, 16 directions, grid={N} nt={NT}", flush=True)
    P = TTIPropagator(grid, m, DX, DY, DZ, DT, device=cuda, zblock=16)
    d_obs = fwi.forward(P, fwi.Shot(src, torch.from_numpy(0.7 * wav), shot.rec_index, shot.rec_ijk), NT).clone()
    J0, g, _ = fwi.gradient(P, shot, d_obs, NT, k=1, snap_mode=1)
    gz = g.sum(dim=(1, 2)).cpu().numpy()
    big = lambda n: torch.tensor([[1.0, 10.0]] * n, dtype=torch.float32)

    def misfit(bz):
        sp = (big(grid.ldx), big(grid.ldy), torch.from_numpy(np.stack([np.ones(grid.ldz), bz], 1).astype(np.float32)))
        Q = TTIPropagator(grid, m, DX, DY, DZ, DT, sponge=sp, device=cuda, zblock=16)
        rec = fwi.forward(Q, shot, NT)
        return 0.5 * float(torch.sum((rec - d_obs).double() ** 2))

    rng = np.random.default_rng(16)
    zz = np.arange(G.nz) / G.nz
    worst = 0.0
    for _ in range(16):
        a, ph = rng.uniform(-1, 1, 3), rng.uniform(0, 2 * np.pi, 3)
        delta = np.round(64 * sum(a[j] * np.cos(np.pi * (j + 1) * zz + ph[j]) for j in range(3))) / 64
        bz = lambda h: np.concatenate([np.ones(G.g), 1.0 + h * delta, np.ones(G.ldz - G.g - G.nz)])
        fd = [(misfit(bz(h)) - misfit(bz(-h))) / (2 * h) for h in (2.0**-7, 2.0**-8)]
        fd = (4 * fd[1] - fd[0]) / 3
        an = float(np.dot(gz, delta))
        err = abs(fd - an) / max(abs(fd), abs(an))
        worst = max(worst, err)
        assert err <= 1e-3, (fd, an, err)
    print(f"worst relative difference over 16 directions: {worst:.2e}", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("mode,k,tol", [(1, 1, 1e-5), (2, 2, 2e-3)])
def test_gpu_density_image_vs_oracle(mode, k, tol, cuda):
    """The density image on the fused path (sigma snapshots beside u, w in the same encoding; tti_rho_image after each
    snapshot step's adjoint step and residual injection, FP64 accumulation) against the FP64 oracle's composition
    (forward derivatives of the stored sigma, then the density-gradient operator); the Vp image of the same run is
    unchanged by storing sigma. A checkpointed run (one level of 30 steps) gives the same bits."""
    import torch

    from miniappsforhpe_amd import fwi
    from miniappsforhpe_amd.grid import Grid
    from miniappsforhpe_amd.propagator import TTIPropagator

    G, m, F, src, idx, ijk, wav, d_obs = setup_oracle()
    print(f"density image mode={mode} k={k} grid={N} nt={NT}", flush=True)
    Jo, go, do, gro = F.gradient(src, wav, idx, d_obs, NT, k=k, rho=True)
    grid = Grid(*N)
    P = TTIPropagator(grid, [torch.from_numpy(a.astype(np.float32)) for a in m], DX, DY, DZ, DT, device=cuda,
                      zblock=16)
    shot = fwi.Shot(src, torch.from_numpy(wav), torch.from_numpy(idx).to(cuda), torch.from_numpy(ijk).to(cuda))
    dob = torch.from_numpy(d_obs.astype(np.float32)).to(cuda)
    J, g, d_syn, grho = fwi.gradient(P, shot, dob, NT, k=k, snap_mode=mode, rho=True)
    J1, g1, _ = fwi.gradient(P, shot, dob, NT, k=k, snap_mode=mode)
    held, rec = fwi._walk(NT, (30,))
    plan = fwi.CheckpointPlan(NT, k, (30,), held, 30 // k, 0, rec)
    Jc, gc, _, _, grc = fwi.gradient_checkpointed(P, shot, dob, NT, k=k, snap_mode=mode, plan=plan, rho=True)
    torch.cuda.synchronize()
    assert J == J1 == Jc and torch.equal(g, g1) and torch.equal(g, gc) and torch.equal(grho, grc)
    e = _rel(grho.cpu().numpy(), gro)
    print(f"density image rel-L2 vs FP64 oracle: {e:.3e}", flush=True)
    assert float(np.abs(gro).max()) > 0 and e <= tol, e


@pytest.mark.gpu
def test_gpu_density_gradient_finite_differences(cuda):
    """The GPU density image (FP32 snapshots, k = 1) against centred finite differences of the GPU misfit along 6
    random smooth density directions through the whole forward time loop (rho enters through b = 1/rho); Richardson
    extrapolation of two step sizes. Bar: 1e-3 relative per direction."""
    import torch

    from miniappsforhpe_amd import fwi
    from miniappsforhpe_amd.grid import Grid
    from miniappsforhpe_amd.propagator import TTIPropagator

    G = O.Geom(*N)
    m = [torch.from_numpy(a.astype(np.float32)) for a in smooth_models(G, 7)]
    src, idx, ijk = shot_geometry(G)
    wav = ricker(NT, DT, 30.0) * 1e6
    grid = Grid(*N)
    shot = fwi.Shot(src, torch.from_numpy(wav), torch.from_numpy(idx).to(cuda), torch.from_numpy(ijk).to(cuda))
    print(f"GPU density finite differences, 6 directions, grid={N} nt={NT}", flush=True)
    P = TTIPropagator(grid, m, DX, DY, DZ, DT, device=cuda, zblock=16)
    d_obs = fwi.forward(P, fwi.Shot(src, torch.from_numpy(0.7 * wav), shot.rec_index, shot.rec_ijk), NT).clone()
    _, _, _, grho = fwi.gradient(P, shot, d_obs, NT, k=1, snap_mode=1, rho=True)
    grho = grho.cpu().numpy()
    rho = 1.0 / m[7].double().numpy()

    def misfit(h, delta):
        mm = list(m)
        mm[7] = torch.from_numpy((1.0 / (rho + h * delta)).astype(np.float32))
        Q = TTIPropagator(grid, mm, DX, DY, DZ, DT, device=cuda, zblock=16)
        rec = fwi.forward(Q, shot, NT)
        return 0.5 * float(torch.sum((rec - d_obs).double() ** 2))

    rng = np.random.default_rng(17)
    z, y, x = np.meshgrid(np.arange(G.ldz), np.arange(G.ldy), np.arange(G.ldx), indexing="ij")
    worst = 0.0
    for _ in range(6):
        a, f, ph = rng.uniform(-1, 1, 3), rng.uniform(0.5, 2.0, (3, 3)), rng.uniform(0, 2 * np.pi, 3)
        delta = 100.0 * sum(a[j] * np.cos(2 * np.pi * (f[j, 0] * x / G.ldx + f[j, 1] * y / G.ldy +
                                                          f[j, 2] * z / G.ldz) + ph[j]) for j in range(3))
        fd = [(misfit(h, delta) - misfit(-h, delta)) / (2 * h) for h in (0.2, 0.1)]
        fd = (4 * fd[1] - fd[0]) / 3
        an = float(np.sum(grho * G.interior(delta)))
        err = abs(fd - an) / max(abs(fd), abs(an))
        worst = max(worst, err)
        assert err <= 1e-3, (fd, an, err)
    print(f"worst relative difference over 6 density directions: {worst:.2e}", flush=True)


class _Box:
    """Grid stand-in for the checkpoint planner (sizes only)."""

    def __init__(self, nx, ny, nz):
        self.nx, self.ny, self.nz = nx, ny, nz

    def nbytes(self):
        g = 4
        ldx = -(-(self.nx + 2 * g) // 32) * 32
        return ldx * (self.ny + 2 * g) * (self.nz + 2 * g) * 4


@pytest.mark.parametrize("nt,k", [(90, 2), (2000, 4), (37, 5), (8, 8)])
def test_checkpoint_plan(nt, k):
    from miniappsforhpe_amd.fwi import _walk, checkpoint_bytes, plan_checkpoints, snapshot_bytes

    g = _Box(96, 80, 64)
    snap, ck = snapshot_bytes(g, 2), checkpoint_bytes(g)
    full = -(-nt // k) * snap
    p = plan_checkpoints(g, nt, k, 2, full)
    assert p.lengths == () and p.checkpoints == 0 and p.recomputed_steps == 0
    prev_rec = 0
    for budget in sorted((full - 1, full // 3, full // 10, 3 * ck + 2 * snap, ck + snap), reverse=True):
        try:
            p = plan_checkpoints(g, nt, k, 2, budget)
        except ValueError:
            continue
        assert p.nbytes <= budget and p.nbytes == p.checkpoints * ck + p.snapshots * snap
        assert (p.checkpoints, p.recomputed_steps) == _walk(nt, p.lengths)
        assert all(L % k == 0 for L in p.lengths) and list(p.lengths) == sorted(p.lengths, reverse=True)
        assert p.recomputed_steps >= prev_rec  # less memory never means less recomputation
        prev_rec = p.recomputed_steps
        # no one-level schedule that fits re-runs fewer steps (brute force over the dry run)
        for L in range(k, k * -(-nt // k), k):
            h, r = _walk(nt, (L,))
            if h * ck + (L // k) * snap <= budget:
                assert r >= p.recomputed_steps
    with pytest.raises(ValueError):
        plan_checkpoints(g, nt, k, 2, snap - 1)


def test_checkpoint_plan_fits_the_large_budget():
    """768 x 768 x 512, 4000 steps, FP16 snapshots every 4 steps under 96 GB (full storage needs 1.25 TB)."""
    from miniappsforhpe_amd.fwi import plan_checkpoints
    from miniappsforhpe_amd.grid import Grid

    p = plan_checkpoints(Grid(768, 768, 512), 4000, 4, 2, 96 * 10**9)
    assert p.nbytes <= 96 * 10**9 and len(p.lengths) == 2
    assert p.recomputed_steps < 2 * 4000


@pytest.mark.parametrize("lengths", [(), (10,), (30, 10)])
def test_checkpoint_executor_frees_its_callbacks_at_once(lengths):
    """run_schedule holds the callbacks (and, in gradient_checkpointed, the snapshot store, checkpoints and the
    re-run propagator behind them) only while it runs: with the cyclic collector off, an object the callbacks refer to
    is gone as soon as the schedule returns. (A recursive closure that keeps itself alive would hold it until the
    collector runs: with several shots in a row that kept every shot's device buffers.)"""
    import gc
    import weakref

    from miniappsforhpe_amd import fwi

    class Buffers:
        pass

    def run():
        held = Buffers()
        ids = iter(range(1000))

        def fwd(a, b, first, stored, save_cb):  # steps a .. b-1, checkpoint hook before each step
            if save_cb is not None and held is not None:
                for j in range(a, b):
                    save_cb(j)

        noop = lambda *a, **k: held and None
        fwi.run_schedule(40, lengths, fwd, noop, noop, lambda: next(ids), noop, noop)
        return weakref.ref(held)

    was = gc.isenabled()
    gc.disable()
    try:
        ref = run()
        assert ref() is None
    finally:
        if was:
            gc.enable()


@pytest.mark.parametrize("lengths", [(), (10,), (44,), (90,), (30, 10), (40, 12), (16, 6), (50, 20)])
def test_checkpoint_executor_schedule(lengths, monkeypatch):
    """The recursive executor on stand-in propagators whose state is the time index: the first forward records
    every step once and in order, restores bring back the right time, every adjoint step n (n = nt..1, in order)
    images the snapshot written by forward step n, the residual exists before the first adjoint step, and the
    re-run steps and checkpoint buffers match the plan's dry run."""
    import torch

    from miniappsforhpe_amd import fwi

    nt, k = 90, 2
    log = {"fwd": 0, "rec": [], "adj": []}

    class Grid1:
        nx = ny = nz = 1
        shape = (1,)

    class Prop:
        def __init__(self, *a, **kw):
            self.grid, self.models, self.dx, self.dy, self.dz, self.dt = Grid1(), None, 1, 1, 1, 1.0
            self.device, self.zblock, self.strip, self.sponge = torch.device("cpu"), 16, 0, None
            self.s = [torch.zeros(1), torch.zeros(1)]
            self.u, self.w = torch.zeros(1), torch.zeros(1)
            self.primed, self.next_adj = False, nt

        @property
        def sigma(self):
            return self.s[0], self.s[1]

        def reset(self):
            for t in self.s + [self.u, self.w]:
                t.zero_()
            self.next_adj = nt

        def prime(self):
            self.primed = True

    class Store:
        def __init__(self, grid, mode, cap, device, sigma=False):
            self.slots = [None] * cap

    def fwd_steps(Q, shot, a, b, rec, store, kk, slot0=0, checkpoint=None):
        for j in range(a, b):
            if checkpoint is not None:
                checkpoint(j)
            assert int(Q.u.item()) == j, (j, Q.u.item())
            for t in (*Q.sigma, Q.u, Q.w):
                t += 1
            log["fwd"] += 1
            if rec is not None:
                log["rec"].append(j)
            if store is not None and (j + 1) % kk == 0:
                store.slots[(j + 1) // kk - 1 - slot0] = j + 1

    def adj_steps(P, shot, res, grad, hi, lo, store, kk, slot0=0, rho=None):
        assert res is not None
        for n in range(hi, lo, -1):
            assert P.next_adj == n
            if n % kk == 0:
                assert store.slots[n // kk - 1 - slot0] == n
            log["adj"].append(n)
            P.next_adj -= 1

    monkeypatch.setattr(fwi, "TTIPropagator", Prop)
    monkeypatch.setattr(fwi, "SnapshotStore", Store)
    monkeypatch.setattr(fwi, "_fwd_steps", fwd_steps)
    monkeypatch.setattr(fwi, "_adj_steps", adj_steps)

    class Shot:
        rec_index = torch.zeros(3)

    held, rec = fwi._walk(nt, lengths)
    plan = fwi.CheckpointPlan(nt, k, lengths, held, 0, 0, rec)
    J, grad, d, _ = fwi.gradient_checkpointed(Prop(), Shot(), torch.zeros((nt, 3)), nt, k, 2, plan=plan)
    assert log["rec"] == list(range(nt))
    assert log["adj"] == list(range(nt, 0, -1))
    assert log["fwd"] == nt + rec


@pytest.mark.gpu
@pytest.mark.parametrize("lengths", [(10,), (44,), (30, 10), (40, 12)])
def test_gpu_checkpointed_gradient_is_bit_identical(lengths, cuda):
    """Recomputation from checkpoints (one level of 10 or 44 steps over 90, two levels of 30/10 and 40/12) reproduces
    the stored run exactly: same records, misfit and FP64 gradient bits (FP16 snapshots, k = 2)."""
    import torch

    from miniappsforhpe_amd import fwi
    from miniappsforhpe_amd.grid import Grid
    from miniappsforhpe_amd.propagator import TTIPropagator

    G, m, F, src, idx, ijk, wav, d_obs = setup_oracle()
    print(f"checkpointed gradient lengths={lengths} grid={N} nt={NT}", flush=True)
    grid = Grid(*N)
    P = TTIPropagator(grid, [torch.from_numpy(a.astype(np.float32)) for a in m], DX, DY, DZ, DT, device=cuda,
                      zblock=16)
    shot = fwi.Shot(src, torch.from_numpy(wav), torch.from_numpy(idx).to(cuda), torch.from_numpy(ijk).to(cuda))
    dob = torch.from_numpy(d_obs.astype(np.float32)).to(cuda)
    J0, g0, d0 = fwi.gradient(P, shot, dob, NT, k=2, snap_mode=2)
    held, rec = fwi._walk(NT, lengths)
    plan = fwi.CheckpointPlan(NT, 2, lengths, held, lengths[-1] // 2, 0, rec)
    J1, g1, d1, _ = fwi.gradient_checkpointed(P, shot, dob, NT, k=2, snap_mode=2, plan=plan)
    torch.cuda.synchronize()
    assert torch.equal(d0, d1)
    assert J0 == J1
    assert torch.equal(g0, g1)


def test_gradient_run_shot_slices():
    from miniappsforhpe_amd import gradient_run

    assert gradient_run.parse(["--shots", "8", "--first-shot", "4", "--count", "4"]).selected == [4, 5, 6, 7]
    assert gradient_run.parse(["--shots", "8", "--first-shot", "5"]).selected == [5, 6, 7]
    assert gradient_run.parse([]).selected == [0]
    for bad in (["--shots", "8", "--first-shot", "6", "--count", "4"], ["--shots", "2", "--first-shot", "2"]):
        with pytest.raises(ValueError):
            gradient_run.parse(bad)


@pytest.mark.gpu
def test_gradient_run_module(cuda):
    """python -m miniappsforhpe_amd.gradient_run on a 64^3 grid, 40 steps, 2 shots, FP16 snapshots every 2 under a
    20 MB snapshot budget (one checkpoint level of 20 steps: 20 steps re-run per shot). The module's own set-up
    (sources, receiver patch, observed data) is rebuilt from what it returns, and every shot's stored-snapshot gradient
    through fwi.gradient must give the same misfits and, summed in the same order, the same FP64 image bits."""
    import torch

    from miniappsforhpe_amd import fwi, gradient_run

    print("gradient_run 64^3, 40 steps, 2 shots, 20 MB", flush=True)
    out, total, pb = gradient_run.run(["--nx", "64", "--ny", "64", "--nz", "64", "--nt", "40", "--k", "2",
                                       "--snap-budget-gb", "0.02", "--freq", "25", "--shots", "2"])
    assert out["plan"] == [20] and out["recomputed_steps"] == 20 and out["steps"] == 200 and out["shots"] == 2
    assert out["plan_bytes"] <= 0.02e9 and out["grad_finite"] and out["grad_max_abs"] > 0 and out["misfit"] > 0
    assert out["estimated_total_device_bytes"] > out["plan_bytes"]
    srcs = {shot.src_index for shot, _ in pb.shots}
    assert len(srcs) == 2
    ref, per_shot = None, []
    for (shot, d_obs), J in zip(pb.shots, out["misfits"]):
        J0, g0, _ = fwi.gradient(pb.prop, shot, d_obs, 40, k=2, snap_mode=2)
        assert J0 == J
        per_shot.append(g0)
        ref = g0 if ref is None else ref + g0
    torch.cuda.synchronize()
    assert torch.equal(ref, total)
    # a slice of the layout (a run split over processes) reproduces that shot of the whole run
    print("gradient_run, shot 1 of the same 2-shot layout alone", flush=True)
    out1, total1, _ = gradient_run.run(["--nx", "64", "--ny", "64", "--nz", "64", "--nt", "40", "--k", "2",
                                        "--snap-budget-gb", "0.02", "--freq", "25", "--shots", "2", "--first-shot",
                                        "1", "--count", "1"])
    assert out1["shot_indices"] == [1] and out1["misfits"] == [out["misfits"][1]]
    assert torch.equal(total1, per_shot[1])
