"""GPU parity of the fused single-pass time steps against the FP64 oracle's step compositions, against the
reference-structure GPU path, and the adjoint dot-product test through the fused kernels.

Tolerances (FP32 arithmetic vs FP64): relative L2 <= 2e-5 after a few time steps; dot-product identity to 1e-6
(FP32 sums over ~1e5 points); FP16 snapshots within 2^-11 of the 64-point segment maximum.
"""
import numpy as np
import pytest
import torch

import oracle as O
from test_parity_type2 import f32, from_gpu, make_models, rel_l2, to_gpu

pytestmark = pytest.mark.gpu

DX, DY, DZ = 10.0, 11.0, 9.0
SHAPES = [(70, 19, 67), (40, 33, 150)]


def setup(shape, rng, cuda, sponge=False, zblock=64):
    from miniappsforhpe_amd.grid import Grid
    from miniappsforhpe_amd.propagator import TTIPropagator

    G, OG = Grid(*shape), O.Geom(*shape)
    m = make_models(G, rng)
    sp = None
    if sponge:
        sp = tuple(rng.uniform(0.9, 1.0, (n, 2)).astype(np.float32) for n in (G.ldx, G.ldy, G.ldz))
    P = TTIPropagator(G, [to_gpu(x, cuda) for x in m], DX, DY, DZ, 4e-4,
                      sponge=None if sp is None else [torch.from_numpy(a) for a in sp], device=cuda, zblock=zblock)
    return G, OG, m, sp, P


def interior_random(G, rng):
    a = np.zeros(G.shape)
    G.interior(a)[...] = rng.standard_normal((G.nz, G.ny, G.nx))
    return f32(a)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sponge", [False, True])
def test_fused_forward_vs_oracle(shape, sponge, rng, cuda):
    G, OG, m, sp, P = setup(shape, rng, cuda, sponge)
    s1, s2, u, w = (interior_random(G, rng) for _ in range(4))
    P.s1[0].copy_(to_gpu(s1, cuda))
    P.s2[0].copy_(to_gpu(s2, cuda))
    P.u.copy_(to_gpu(u, cuda))
    P.w.copy_(to_gpu(w, cuda))
    src = int(np.ravel_multi_index((G.ghost + G.nz // 2, G.ghost + G.ny // 2, G.ghost + G.nx // 3), G.shape))
    amps = [0.5, -1.0, 2.0, 0.0, 1.5]
    dt = 4e-4
    P.prime()
    for a in amps:
        P.step_fwd_fused(src, dt * a)
    # oracle
    o1, o2, ou, ow = s1.copy(), s2.copy(), u.copy(), w.copy()
    d = [OG.zeros() for _ in range(6)]
    spo = sp if sp is not None else O.unit_sponge(OG)
    for a in amps:
        O.step_fwd(OG, o1, o2, ou, ow, m, d, spo, O.f32(dt), DX, DY, DZ)
        ou.ravel()[src] += O.f32(dt * a)
        ow.ravel()[src] += O.f32(dt * a)
    g1, g2 = P.sigma_prev
    for got, ref in ((g1, o1), (g2, o2), (P.u, ou), (P.w, ow)):
        e = rel_l2(from_gpu(got), ref)
        assert e < 2e-5, e


@pytest.mark.parametrize("shape", SHAPES[:1])
def test_fused_forward_vs_plain_path(shape, rng, cuda):
    G, OG, m, sp, P = setup(shape, rng, cuda)
    s1, s2, u, w = (interior_random(G, rng) for _ in range(4))
    state = [to_gpu(x, cuda) for x in (s1, s2, u, w)]
    P.s1[0].copy_(state[0]); P.s2[0].copy_(state[1]); P.u.copy_(state[2]); P.w.copy_(state[3])
    P.prime()
    for _ in range(4):
        P.step_fwd_fused()
    fused = [from_gpu(t) for t in (*P.sigma_prev, P.u, P.w)]
    P.reset()
    P.s1[0].copy_(state[0]); P.s2[0].copy_(state[1]); P.u.copy_(state[2]); P.w.copy_(state[3])
    for _ in range(4):
        P.step_fwd_plain()
    plain = [from_gpu(t) for t in (*P.sigma, P.u, P.w)]
    for a, b in zip(fused, plain):
        assert rel_l2(a, b) < 1e-5


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("sponge", [False, True])
def test_fused_adjoint_vs_oracle(shape, sponge, rng, cuda):
    G, OG, m, sp, P = setup(shape, rng, cuda, sponge)
    a1, a2, u, w = (interior_random(G, rng) for _ in range(4))
    P.s1[0].copy_(to_gpu(a1, cuda)); P.s2[0].copy_(to_gpu(a2, cuda))
    P.u.copy_(to_gpu(u, cuda)); P.w.copy_(to_gpu(w, cuda))
    for _ in range(4):
        P.step_adj_fused()
    o1, o2, ou, ow = a1.copy(), a2.copy(), u.copy(), w.copy()
    d8 = [OG.zeros() for _ in range(8)]
    spo = sp if sp is not None else O.unit_sponge(OG)
    for _ in range(4):
        O.step_adj(OG, o1, o2, ou, ow, m, d8, spo, O.f32(4e-4), DX, DY, DZ, sponge_active=sp is not None)
    g1, g2 = P.sigma
    for got, ref in ((g1, o1), (g2, o2), (P.u, ou), (P.w, ow)):
        e = rel_l2(from_gpu(got), ref)
        assert e < 2e-5, e


def test_fused_adjoint_vs_plain_path(rng, cuda):
    G, OG, m, sp, P = setup((70, 19, 67), rng, cuda, sponge=True)
    a1, a2, u, w = (to_gpu(interior_random(G, rng), cuda) for _ in range(4))
    P.s1[0].copy_(a1); P.s2[0].copy_(a2); P.u.copy_(u); P.w.copy_(w)
    for _ in range(3):
        P.step_adj_fused()
    fused = [from_gpu(t) for t in (*P.sigma, P.u, P.w)]
    P.reset()
    P.s1[0].copy_(a1); P.s2[0].copy_(a2); P.u.copy_(u); P.w.copy_(w)
    for _ in range(3):
        P.step_adj_plain()
    plain = [from_gpu(t) for t in (*P.sigma, P.u, P.w)]
    for a, b in zip(fused, plain):
        assert rel_l2(a, b) < 1e-5


@pytest.mark.parametrize("shape,zblock", [(SHAPES[0], 64), (SHAPES[1], 16)])
def test_fused_free_surface_vs_oracle(shape, zblock, rng, cuda):
    """Stress-free surface at the top interior plane (the reference's mirrors of sigma and of the derivatives):
    5 fused forward steps, then 4 fused adjoint steps, each against the oracle's step composition with
    free_surface=True (z chunks of 16 planes in the second case, so the mirrored plane lies in the first of many)."""
    G, OG, m, sp, P = setup(shape, rng, cuda, zblock=zblock)
    P.free_surface = True
    s1, s2, u, w = (interior_random(G, rng) for _ in range(4))
    P.s1[0].copy_(to_gpu(s1, cuda)); P.s2[0].copy_(to_gpu(s2, cuda))
    P.u.copy_(to_gpu(u, cuda)); P.w.copy_(to_gpu(w, cuda))
    dt = 4e-4
    P.prime()
    for _ in range(5):
        P.step_fwd_fused()
    o1, o2, ou, ow = s1.copy(), s2.copy(), u.copy(), w.copy()
    d = [OG.zeros() for _ in range(6)]
    spo = O.unit_sponge(OG)
    for _ in range(5):
        O.step_fwd(OG, o1, o2, ou, ow, m, d, spo, O.f32(dt), DX, DY, DZ, free_surface=True)
    g1, g2 = P.sigma_prev
    errs = [rel_l2(from_gpu(got), ref) for got, ref in ((g1, o1), (g2, o2), (P.u, ou), (P.w, ow))]
    print(f"free surface, forward: rel L2 {max(errs):.2e}", flush=True)
    assert max(errs) < 2e-5, errs
    # the mirrored planes of sigma: zero at izbeg, negated images above
    top = from_gpu(g1)[G.ghost - 4: G.ghost + 5]
    assert np.all(top[4] == 0) and np.array_equal(top[:4], -top[8:4:-1])
    # adjoint from the oracle's state (its sigma mirrored as the forward left it)
    a1, a2, au, aw = o1.copy(), o2.copy(), ou.copy(), ow.copy()
    P.reset()
    P.s1[0].copy_(to_gpu(a1, cuda)); P.s2[0].copy_(to_gpu(a2, cuda))
    P.u.copy_(to_gpu(au, cuda)); P.w.copy_(to_gpu(aw, cuda))
    for _ in range(4):
        P.step_adj_fused()
    d8 = [OG.zeros() for _ in range(8)]
    for _ in range(4):
        O.step_adj(OG, a1, a2, au, aw, m, d8, spo, O.f32(dt), DX, DY, DZ, free_surface=True)
    g1, g2 = P.sigma
    errs = [rel_l2(from_gpu(got), ref) for got, ref in ((g1, a1), (g2, a2), (P.u, au), (P.w, aw))]
    print(f"free surface, adjoint: rel L2 {max(errs):.2e}", flush=True)
    assert max(errs) < 2e-5, errs


def test_fused_free_surface_vs_plain_path(rng, cuda):
    """With the free surface, the fused steps against the reference-structure GPU steps (mirrors as separate
    kernels): 4 forward, then 3 adjoint steps."""
    G, OG, m, sp, P = setup((70, 19, 67), rng, cuda)
    P.free_surface = True
    state = [to_gpu(interior_random(G, rng), cuda) for _ in range(4)]

    def load():
        P.reset()
        P.s1[0].copy_(state[0]); P.s2[0].copy_(state[1]); P.u.copy_(state[2]); P.w.copy_(state[3])

    load()
    P.prime()
    for _ in range(4):
        P.step_fwd_fused()
    mid = [t.clone() for t in (*P.sigma, P.u, P.w)]  # sigma for the next step, already time-updated and mirrored
    fused_f = [from_gpu(t) for t in (*P.sigma_prev, P.u, P.w)]
    for _ in range(3):
        P.step_adj_fused()
    fused_a = [from_gpu(t) for t in (*P.sigma, P.u, P.w)]
    load()
    for _ in range(4):
        P.step_fwd_plain()
    plain_f = [from_gpu(t) for t in (*P.sigma, P.u, P.w)]
    P.reset()
    for t, v in zip((P.s1[0], P.s2[0], P.u, P.w), mid):
        t.copy_(v)
    for _ in range(3):
        P.step_adj_plain()
    plain_a = [from_gpu(t) for t in (*P.sigma, P.u, P.w)]
    for a, b in zip(fused_f + fused_a, plain_f + plain_a):
        assert rel_l2(a, b) < 1e-5


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_fused_imaging_condition(mode, rng, cuda):
    """grad (FP64) += u' * u_snap + w' * w_snap inside the fused adjoint, against the oracle on the decoded snapshot."""
    from miniappsforhpe_amd.propagator import SnapshotStore

    G, OG, m, sp, P = setup((70, 19, 67), rng, cuda)
    store = SnapshotStore(G, mode, 1, cuda)
    s1, s2, u, w = (interior_random(G, rng) for _ in range(4))
    P.s1[0].copy_(to_gpu(s1, cuda)); P.s2[0].copy_(to_gpu(s2, cuda))
    P.u.copy_(to_gpu(u, cuda)); P.w.copy_(to_gpu(w, cuda))
    P.prime()
    P.step_fwd_fused(snap=store.desc(0))
    uf, wf = from_gpu(G.interior(P.u)), from_gpu(G.interior(P.w))
    du, dw = from_gpu(store.decode(0, 3)), from_gpu(store.decode(0, 4))
    tol = {1: 0.0, 2: 2.0**-11, 3: 2.0**-8}[mode]
    # per 64-point segment, |decoded - exact| <= tol * segment max (FP16: power-of-two scale, 11-bit mantissa)
    for ex, de in ((uf, du), (wf, dw)):
        if mode == 1:
            assert np.array_equal(de, ex)
        else:
            assert np.abs(de - ex).max() <= tol * np.abs(ex).max()
    # adjoint step with imaging
    a1, a2, au, aw = (interior_random(G, rng) for _ in range(4))
    P.s1[P.cur].copy_(to_gpu(a1, cuda)); P.s2[P.cur].copy_(to_gpu(a2, cuda))
    P.u.copy_(to_gpu(au, cuda)); P.w.copy_(to_gpu(aw, cuda))
    grad = torch.zeros((G.nz, G.ny, G.nx), dtype=torch.float64, device=cuda)
    P.step_adj_fused(snap=store.desc(0), grad=grad)
    ua, wa = from_gpu(G.interior(P.u)), from_gpu(G.interior(P.w))
    ref = ua * du + wa * dw
    assert rel_l2(grad.cpu().numpy(), ref) < 1e-6


@pytest.mark.parametrize("mode", [2, 3])
def test_snapshot_encoding_is_exact(mode, rng, cuda):
    """The snapshot written by the fused forward equals, bit for bit, the host encoding of the step's u and w: FP16
    halves of v * 2^(14 - e) with e the binary exponent of the 64-point segment maximum (max < 2^e) and the scale
    2^(e - 14); BF16 round-to-nearest-even. Rows mix magnitudes over 12 decades and the x extent ends in a partial
    segment, so the wave-wide maximum and the segment bookkeeping are both exercised."""
    from miniappsforhpe_amd.propagator import SnapshotStore

    G, OG, m, sp, P = setup((70, 19, 67), rng, cuda)
    store = SnapshotStore(G, mode, 1, cuda)
    vals = []
    for _ in range(4):
        a = interior_random(G, rng)
        mag = 10.0 ** rng.uniform(-6, 6, (G.ldz, G.ldy, 1))
        vals.append(f32(a * mag))
    P.s1[0].copy_(to_gpu(vals[0], cuda)); P.s2[0].copy_(to_gpu(vals[1], cuda))
    P.u.copy_(to_gpu(vals[2], cuda)); P.w.copy_(to_gpu(vals[3], cuda))
    P.prime()
    P.step_fwd_fused(snap=store.desc(0))
    torch.cuda.synchronize()
    for field, arr, sc in ((P.u, store.u, store.su), (P.w, store.w, store.sw)):
        v = G.interior(from_gpu(field)).astype(np.float32)  # (nz, ny, nx)
        got = arr[0].cpu()[:, :, : G.nx]
        if mode == 3:
            ref = torch.from_numpy(v).to(torch.bfloat16)
            assert torch.equal(got.view(torch.int16), ref.view(torch.int16))
            continue
        nseg = -(-G.nx // 64)
        pad = np.zeros((G.nz, G.ny, nseg * 64), dtype=np.float32)
        pad[:, :, : G.nx] = v
        mx = np.abs(pad.reshape(G.nz, G.ny, nseg, 64)).max(axis=3)
        e = np.frexp(mx)[1].astype(np.int64)
        e[mx == 0] = 0
        scale = np.ldexp(np.float32(1.0), e - 14).astype(np.float32)
        assert np.array_equal(sc[0].cpu().numpy()[:, :, :nseg], scale)
        ex = np.repeat(e, 64, axis=2)[:, :, : G.nx]
        half = np.ldexp(v, 14 - ex).astype(np.float32).astype(np.float16)
        assert np.array_equal(got.numpy().view(np.uint16), half.view(np.uint16))


@pytest.mark.parametrize("zblock,strip", [(16, 3), (40, 1), (200, 2)])
def test_block_schedule_keeps_results(zblock, strip, rng, cuda):
    """The z depth of the blocks and the strip schedule of the tiles (4 x tiles here: strips of 3 + 1, 1, 2 + 2)
    change which block computes a point, never its value: forward steps with an FP16 snapshot and adjoint steps with
    FP64 imaging are bit-identical to the default schedule."""
    from miniappsforhpe_amd.propagator import SnapshotStore

    G, OG, m, sp, P = setup((200, 21, 37), rng, cuda)
    init = [to_gpu(interior_random(G, rng), cuda) for _ in range(4)]

    def run(zb, st):
        P.reset()
        P.zblock, P.strip = zb, st
        P.s1[0].copy_(init[0]); P.s2[0].copy_(init[1]); P.u.copy_(init[2]); P.w.copy_(init[3])
        store = SnapshotStore(G, 2, 1, cuda)
        grad = torch.zeros((G.nz, G.ny, G.nx), dtype=torch.float64, device=cuda)
        P.prime()
        for i in range(3):
            P.step_fwd_fused(snap=store.desc(0) if i == 2 else None)
        for i in range(2):
            P.step_adj_fused(snap=store.desc(0), grad=grad)
        torch.cuda.synchronize()
        return [t.clone() for t in (*P.sigma, P.u, P.w, store.u, store.su, grad)]

    ref = run(64, 0)
    got = run(zblock, strip)
    for a, b in zip(ref, got):
        assert torch.equal(a, b)


@pytest.mark.parametrize("nsteps", [1, 8])
def test_adjoint_dot_product_fused(rng, cuda, nsteps):
    """<P x, y> = <x, P^T y> for nsteps fused forward steps vs nsteps fused adjoint steps (no sponge, zero ghosts),
    FP32 kernels, dot products summed in FP64. The relative defect is printed; the bar, 1e-6, is a few FP32 roundings
    (2^-23 = 1.2e-7) of the operator's values, far below any error of a wrong coefficient, index or sign."""
    G, OG, m, sp, P = setup((48, 40, 44), rng, cuda)
    x = [to_gpu(interior_random(G, rng), cuda) for _ in range(4)]
    y = [to_gpu(interior_random(G, rng), cuda) for _ in range(4)]
    P.s1[0].copy_(x[0]); P.s2[0].copy_(x[1]); P.u.copy_(x[2]); P.w.copy_(x[3])
    P.prime()
    for _ in range(nsteps):
        P.step_fwd_fused()
    fs1, fs2 = P.sigma_prev
    dot = lambda a, b: float(torch.sum(a.double() * b.double()))
    lhs = dot(fs1, y[0]) + dot(fs2, y[1]) + dot(P.u, y[2]) + dot(P.w, y[3])
    P.reset()
    P.s1[0].copy_(y[2]); P.s2[0].copy_(y[3]); P.u.copy_(y[0]); P.w.copy_(y[1])
    for _ in range(nsteps):
        P.step_adj_fused()
    as1, as2 = P.sigma
    rhs = dot(x[0], P.u) + dot(x[1], P.w) + dot(x[2], as1) + dot(x[3], as2)
    defect = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"fused dot-product test, {nsteps} steps: <Px,y> = {lhs:.9e}, <x,P^T y> = {rhs:.9e}, rel defect {defect:.2e}")
    assert defect <= 1e-6, (lhs, rhs, defect)


def test_gather_and_inject(rng, cuda):
    from miniappsforhpe_amd._lib import call
    from miniappsforhpe_amd.grid import Grid

    G = Grid(40, 30, 20)
    s1, s2 = (to_gpu(interior_random(G, rng), cuda) for _ in range(2))
    pts = [(G.ghost + 5, G.ghost + 7, G.ghost + 3), (G.ghost + 39, G.ghost + 29, G.ghost + 19)]
    idx = torch.tensor([np.ravel_multi_index((z, y, x), G.shape) for x, y, z in pts], dtype=torch.int64, device=cuda)
    ijk = torch.tensor([c for p in pts for c in p], dtype=torch.int32, device=cuda)
    out = torch.zeros(2, dtype=torch.float32, device=cuda)
    call("tti_gather_pressure", s1.data_ptr(), s2.data_ptr(), idx.data_ptr(), 2, out.data_ptr(),
         torch.cuda.current_stream().cuda_stream)
    ref = -(2 * s1.view(-1)[idx] + s2.view(-1)[idx]) / 3
    assert torch.allclose(out, ref, rtol=1e-6)
    u, w, o1, o2 = (torch.zeros(G.shape, device=cuda) for _ in range(4))
    res = torch.tensor([3.0, -6.0], device=cuda)
    call("tti_inject_residual", u.data_ptr(), w.data_ptr(), o1.data_ptr(), o2.data_ptr(), idx.data_ptr(),
         ijk.data_ptr(), 2, res.data_ptr(), 0.5, None, None, None, *G.ranges, None, None,
         torch.cuda.current_stream().cuda_stream)
    assert torch.allclose(u.view(-1)[idx], torch.tensor([-2.0, 4.0], device=cuda))
    assert torch.allclose(w.view(-1)[idx], torch.tensor([-1.0, 2.0], device=cuda))
    assert torch.allclose(o1.view(-1)[idx], torch.tensor([-1.0, 2.0], device=cuda))
    assert float(u.abs().sum()) == pytest.approx(6.0)
