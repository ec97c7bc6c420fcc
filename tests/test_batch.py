"""Batched multi-shot steps (include/tti_batch.h, miniappsforhpe_amd.batch).

CPU: the ctypes mirror of tti_shot has the C layout; aperture model cutting takes the right window; the batch
constructor rejects inconsistent shot tables before touching the GPU.

GPU: three shots with uneven apertures cut from one model run as a batch (one launch per step for all shots) and
reproduce, bit for bit, each shot run alone through fwi.gradient (records, misfit, FP64 image; FP16 snapshots every 2
steps, z blocks of 16 planes, strips of 2 x tiles) — the single-shot path is itself pinned against the FP64 oracle
in test_gradient.py.
hipGraph replays of the captured passes, with eager work and synchronisations between them, equal the eager passes.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from fwi_oracle import ricker, smooth_models

MODEL_N = (160, 40, 36)
NT, K, DT, DX = 40, 2, 1.0e-3, 10.0


def apertures():
    from miniappsforhpe_amd.batch import Aperture

    out = []
    for (x0, y0, z0, nx, ny, nz) in [(0, 0, 0, 28, 24, 26), (12, 8, 6, 136, 20, 30), (20, 16, 10, 20, 24, 24)]:
        sx, sy, sz = x0 + nx // 3, y0 + ny // 2, z0 + nz // 3
        rec = np.array([(sx - 4 + j, sy + (j % 3) - 1, sz + 6) for j in range(10)])
        out.append(Aperture(x0, y0, z0, nx, ny, nz, (sx, sy, sz), rec))
    return out


def test_shot_struct_layout_matches_c():
    from miniappsforhpe_amd._lib import TTIShot, lib

    assert lib().tti_batch_shot_bytes() == C.sizeof(TTIShot)


def test_cut_models_window():
    import torch

    from miniappsforhpe_amd.batch import cut_models
    from miniappsforhpe_amd.grid import Grid

    G = Grid(*MODEL_N)
    vols = [torch.arange(G.ntot, dtype=torch.float32).reshape(G.shape) + 1e6 * i for i in range(8)]
    for ap in apertures():
        sg = ap.grid()
        cut = cut_models(vols, G, ap)
        for i, c in enumerate(cut):
            assert tuple(c.shape) == sg.shape
            # interior + ghost rim: the model's array at offset (x0, y0, z0) (both ghosts are 4 points)
            ref = vols[i][ap.z0: ap.z0 + sg.ldz, ap.y0: ap.y0 + sg.ldy, ap.x0: ap.x0 + ap.nx + 8]
            assert torch.equal(c[:, :, : ap.nx + 8], ref)
            # x padding replicates the last column of the window
            assert torch.equal(c[:, :, ap.nx + 8:], ref[:, :, -1:].expand(-1, -1, sg.ldx - ap.nx - 8))


def test_batch_create_rejects_bad_tables():
    from miniappsforhpe_amd._lib import TTIShot, lib

    h = C.c_void_p()
    assert lib().tti_batch_create(None, 0, 1e-3, 10.0, 10.0, 10.0, C.byref(h)) == -1
    s = (TTIShot * 1)()
    assert lib().tti_batch_create(s, 1, 1e-3, 10.0, 10.0, 10.0, C.byref(h)) == -1  # null fields
    assert h.value is None


def _setup(cuda):
    import torch

    from miniappsforhpe_amd.batch import BatchFWI
    from miniappsforhpe_amd.grid import Grid

    G = Grid(*MODEL_N)
    m = smooth_models(O.Geom(*MODEL_N), 5)
    models = [torch.from_numpy(a.astype(np.float32)).to(cuda) for a in m]
    wav = torch.from_numpy(ricker(NT, DT, 30.0) * 1e6)
    # z blocks of 16 planes and strips of 2 x tiles (the 136-wide aperture has 3): the single-shot path below runs the
    # default schedule, so agreement also pins the batched tile decode
    B = BatchFWI(models, G, apertures(), wav, DX, DX, DX, DT, NT, K, snap_mode=2, zblock=16, strip=2, device=cuda)
    return B, wav


@pytest.mark.gpu
def test_gpu_batch_matches_single_shots(cuda):
    import torch

    from miniappsforhpe_amd import fwi
    from miniappsforhpe_amd.propagator import TTIPropagator

    print(f"batched gradient: 3 shots from {MODEL_N}, nt={NT} k={K}", flush=True)
    B, wav = _setup(cuda)
    B.forward(use_graph=False)
    d_obs = [0.7 * s.records[:, : s.nrec].clone() for s in B.shots]
    J, grads, total = B.gradient(d_obs, use_graph=False)
    torch.cuda.synchronize()
    ref_total = torch.zeros_like(total)
    for i, s in enumerate(B.shots):
        print(f"shot {i}: aperture {s.grid.nx}x{s.grid.ny}x{s.grid.nz}", flush=True)
        P = TTIPropagator(s.grid, s.models, DX, DX, DX, DT, device=cuda, zblock=16)
        shot = fwi.Shot(s.src_index, wav, s.rec_index, s.rec_ijk)
        J1, g1, d1 = fwi.gradient(P, shot, d_obs[i], NT, k=K, snap_mode=2)
        torch.cuda.synchronize()
        assert torch.equal(d1, s.records[:, : s.nrec])
        assert J1 == J[i]
        assert torch.equal(g1, grads[i])
        ap = s.ap
        ref_total[ap.z0: ap.z0 + ap.nz, ap.y0: ap.y0 + ap.ny, ap.x0: ap.x0 + ap.nx] += g1
        assert float(g1.abs().max()) > 0
    assert torch.equal(ref_total, total)
    B.close()


@pytest.mark.gpu
def test_gpu_batch_graph_replays(cuda):
    import torch

    print("batched passes: hipGraph replays vs eager", flush=True)
    B, _ = _setup(cuda)
    B.forward(use_graph=False)
    rec0 = [s.records.clone() for s in B.shots]
    d_obs = [0.5 * r[:, : s.nrec] for r, s in zip(rec0, B.shots)]
    J0, g0, t0 = B.gradient(d_obs, use_graph=False)
    g0 = [g.clone() for g in g0]
    t0 = t0.clone()
    for rep in range(3):
        x = torch.randn(256, 256, device=cuda)
        (x @ x).sum().item()  # eager work and a synchronisation between replays
        for s in B.shots:
            s.records.fill_(7.0)
        J, g, t = B.gradient(d_obs, use_graph=True)
        torch.cuda.synchronize()
        for r, s in zip(rec0, B.shots):
            assert torch.equal(r, s.records), rep
        assert J == J0, rep
        for a, b in zip(g0, g):
            assert torch.equal(a, b), rep
        assert torch.equal(t0, t), rep
    B.close()
