"""The batched multi-shot FWI gradient at its workload size: shots with seeded uneven apertures cut from one model, all
advancing together (one launch per step for every shot), FP16 snapshots under a shared checkpoint schedule, per-shot
FP64 images summed into the model-sized gradient in FP64.

    python -m miniappsforhpe_amd.batch_run --shots 64 --nt 1500 --k 4 --snap-budget-gb 150

Defaults are the workload: 64 apertures of 128-320 cells per axis (seed 1004) from a 1024 x 1024 x 400 layered TTI
model, 1500 steps, 12 Hz Ricker sources 8 planes below each aperture's top, 64 receivers per shot, observed data =
0.8 x each shot's synthetic data (made beforehand by a forward pass without snapshots). The block schedule (z depth,
strip width) is chosen by the autotuner. `--snap-budget-gb` bounds the checkpoints plus snapshot slots of all shots
together; the rest of the device footprint (models, two state sets, images) is fixed by the apertures.
Prints one JSON line: plan, timings of the gradient (first forward, re-run forwards, adjoint) and updates/s.
"""
from __future__ import annotations

import argparse
import json
import time

import torch

from . import autotune
from .batch import BatchFWI, seeded_apertures
from .grid import Grid
from .models import cfl_dt, layered_tti, ricker
from .propagator import SNAP_FP16


def main(argv=None) -> dict:
    ap = argparse.ArgumentParser()
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--nt", type=int, default=1500)
    ap.add_argument("--k", type=int, default=4)
    # 150 GB of checkpoints and slots leave the 64-shot run at a 224 GB device peak (models, two state sets and
    # images take the rest), clear of the 288 GB of one MI355X
    ap.add_argument("--snap-budget-gb", type=float, default=150.0)
    ap.add_argument("--freq", type=float, default=12.0)
    ap.add_argument("--seed", type=int, default=1004)
    ap.add_argument("--model", type=int, nargs=3, default=(1024, 1024, 400))
    args = ap.parse_args(argv)

    dev = torch.device("cuda", 0)
    MG = Grid(*args.model)
    models, vmax = layered_tti(MG, seed=args.seed, device=dev)
    dt = cfl_dt(vmax, 10.0, 10.0, 10.0)
    aps = seeded_apertures(MG, args.shots, seed=args.seed)
    wav = ricker(args.nt, dt, args.freq) * 1e4
    # sizes for the planner from a batch with one snapshot slot per shot, then the batch with the plan's slots
    B = BatchFWI(models, MG, aps, wav, 10.0, 10.0, 10.0, dt, args.nt, args.k, snap_mode=SNAP_FP16, device=dev,
                 slots=1)
    plan = B.plan(int(args.snap_budget_gb * 1e9))
    B.close()
    del B
    torch.cuda.empty_cache()
    B = BatchFWI(models, MG, aps, wav, 10.0, 10.0, 10.0, dt, args.nt, args.k, snap_mode=SNAP_FP16, device=dev,
                 slots=plan.snapshots)
    del models
    torch.cuda.empty_cache()
    tuned = autotune.tune_batch(B)
    print(f"{len(aps)} shots, {B.interior_points()} interior points; plan {plan.lengths}: {plan.checkpoints} "
          f"checkpoints + {plan.snapshots} snapshot slots = {plan.nbytes / 1e9:.1f} GB, {plan.recomputed_steps} steps "
          f"re-run; schedule zblock {tuned['zblock']} strip {tuned['strip']}", flush=True)
    t0 = time.perf_counter()
    d_obs = [0.8 * r for r in B.forward_records()]
    torch.cuda.synchronize()
    t_obs = time.perf_counter() - t0
    print(f"observed data: {t_obs:.1f} s", flush=True)
    torch.cuda.reset_peak_memory_stats(dev)
    t0 = time.perf_counter()
    def progress(lo, hi):
        print(f"adjoint steps {hi} .. {lo + 1} done, {time.perf_counter() - t0:.1f} s", flush=True)

    J, grads, total = B.gradient_checkpointed(d_obs, plan, progress)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    pts = B.interior_points()
    steps = 2 * args.nt + plan.recomputed_steps
    out = {
        "workload": f"batched FWI gradient, {len(aps)} shots (apertures 128-320 per axis, seed {args.seed}) from a "
                    f"{MG.nx}x{MG.ny}x{MG.nz} model, {args.nt} steps, FP16 snapshots every {args.k} under "
                    f"{args.snap_budget_gb:g} GB of checkpoints and snapshot slots, FP64 images summed",
        "interior_points": pts, "plan": list(plan.lengths), "checkpoints": plan.checkpoints,
        "snapshot_slots": plan.snapshots, "plan_bytes": plan.nbytes, "recomputed_steps": plan.recomputed_steps,
        "zblock": tuned["zblock"], "strip": tuned["strip"], "autotune_ms": tuned["ms"],
        "observed_data_s": t_obs, "seconds": secs, "steps": steps,
        "updates_per_s": pts * steps / secs,
        "misfit": float(sum(J)), "grad_max_abs": float(total.abs().max()),
        "grad_finite": bool(torch.isfinite(total).all()),
        "peak_device_bytes": torch.cuda.max_memory_allocated(dev),
    }
    B.close()
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
