"""Autotuner for the fused time steps: picks the z depth of the thread blocks (planes per block, the length of the
register queue's march) and the block schedule (x tiles per strip) by timing candidate launches on the caller's own
volumes.

The tile shape in x-y is fixed by the kernels (64 x 4 rows forward, 64 x 8 rows adjoint, two and one blocks per CU).
The z depth trades the 8 halo planes every block re-reads and recomputes against the number of blocks that fill the
256 CUs: a 512^3 volume runs best at 256 planes, a batch of small apertures or a single small volume may want fewer.
The strip width decides which tiles run at the same time on one XCD: full rows share x halos perfectly, narrower
strips bring more y neighbours into the same L2 at once. Timing is by HIP events on the current stream around a few
forward and adjoint steps (no snapshots, no imaging: those add the same cost to every candidate). Neither knob
changes a result bit (each point is computed by the same code whichever block computes it).
"""
from __future__ import annotations

import itertools

import torch

ZBLOCKS = (64, 128, 256)
STRIPS = (0, 2, 4)


def _time(fn, reps: int) -> float:
    fn()  # warm (and the first launch of each instantiation)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _candidates(nz: int, zblocks, strips):
    return sorted({(min(z, nz), s) for z, s in itertools.product(zblocks, strips)})


def tune_propagator(P, zblocks=ZBLOCKS, strips=STRIPS, reps: int = 3) -> dict:
    """Sets P.zblock and P.strip to the fastest candidate for one fused forward + one fused adjoint step on P's grid.
    P is reset afterwards. Returns {"zblock", "strip", "ms": {"zblock/strip": ms per forward+adjoint pair}}."""
    ms = {}
    P.reset()
    P.prime()
    for zb, st in _candidates(P.grid.nz, zblocks, strips):
        P.zblock, P.strip = zb, st
        ms[(zb, st)] = _time(lambda: (P.step_fwd_fused(), P.step_adj_fused()), reps)
    best = min(ms, key=ms.get)
    P.zblock, P.strip = best
    P.reset()
    return {"zblock": best[0], "strip": best[1], "ms": {f"{z}/{s}": round(v, 4) for (z, s), v in ms.items()}}


def tune_batch(B, zblocks=ZBLOCKS, strips=STRIPS, reps: int = 2) -> dict:
    """Rebuilds the batch with the fastest (z depth, strip) for one batched forward + adjoint step (all shots in one
    launch each). The shots' state is left at rest. Returns {"zblock", "strip", "ms": {...}}."""
    from ._lib import call

    def reset():
        call("tti_batch_reset", B._h, torch.cuda.current_stream().cuda_stream)

    ms = {}
    for zb, st in _candidates(max(s.grid.nz for s in B.shots), zblocks, strips):
        B.retile(zb, st)
        reset()
        ms[(zb, st)] = _time(lambda: (B.step_fwd(0, 0), B.step_adj(1, 1)), reps)
    best = min(ms, key=ms.get)
    B.retile(*best)
    reset()
    return {"zblock": best[0], "strip": best[1], "ms": {f"{z}/{s}": round(v, 4) for (z, s), v in ms.items()}}
