#!/usr/bin/env python3
"""Headline benchmark: TTI 8th-order forward + adjoint (with zero-lag imaging) at 512^3 on one MI355X.

One bench "step" = one shot's RTM/FWI gradient pass over the synthetic 512^3 TTI subdomain:
    nt fused forward steps (FP16 snapshots of u, w every `k` steps)  +  nt fused adjoint steps (imaging condition
    g += u' u_snap + w' w_snap in FP64 at the snapshot steps).
Grid-point updates per step = 512^3 * 2 nt. `value` = updates/s of the whole job with all inputs resident in HBM.

The line also reports
  * `plain_fp32`: the same fwd+adj pass run through the reference-structure kernels (3 kernels per forward step,
    3 per adjoint step, FP32 snapshots, FP32 imaging kernel) and the fused/plain speedup;
  * `roofline`: algorithmic HBM bytes of the dominant fused kernel instantiation (the one with the largest share
    of the pass; 64 B per grid-point update: read sigma x2, u, w, 8 model words; write sigma x2, u, w) / its mean
    duration from HIP events on the launch stream, vs 8 TB/s; `traffic` = its HBM bytes per launch, measured in
    this run by two child rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) over a short run of the same workload
    and schedule, each scaled by what the counter reads for a 1 GiB device-to-device copy (falls back to the
    committed counter summary of the workload, named in `traffic_source`, when rocprofv3 is absent or fails);
  * `cpu_baseline`: the FP32 OpenMP CPU restatement of the same fwd+adj method on this host's cores (a bounded
    sample, see its `sample` field);
  * `etti_reference_cpu`: the reference's own CPU driver (cpuMiniApp/simulator4, staggered elastic TTI, built into
    oracle/_ref) on 8 MPI ranks beside this library's GPU simulator3SetUp on the same zone layout.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = 64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512, help="interior points per axis")
    ap.add_argument("--nt", type=int, default=100, help="forward (and adjoint) time steps per bench step")
    ap.add_argument("--k", type=int, default=4, help="snapshot / imaging interval (xcorr_step)")
    ap.add_argument("--snap", choices=["fp16", "bf16", "fp32"], default="fp16")
    ap.add_argument("--zblock", type=int, default=0, help="z planes per block (0: autotuned with the strip)")
    ap.add_argument("--strip", type=int, default=0, help="x tiles per strip of the block schedule (0: full rows)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-shot measurement")
    ap.add_argument("--batch-shots", type=int, default=64)
    ap.add_argument("--batch-nt", type=int, default=12, help="forward (and adjoint) steps of the batched measurement")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run one 1 GiB device-to-device copy first (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the live rocprofv3 traffic passes (set automatically under a profiler)")
    return ap.parse_args()


# template arguments of the fused instantiation behind each timed launch kind (snapshot mode numbers as tti_fused.h)
SNAP_NUM = {"fp32": 1, "fp16": 2, "bf16": 3}


def kernel_pattern(kind: str, snap: str) -> str:
    m = SNAP_NUM[snap]
    return {"fwd": "k_fused_fwd<4, 0, false,", "fwd+snapshot": f"k_fused_fwd<4, {m}, false,",
            "adj": "k_fused_adj<0, false>", "adj+imaging": f"k_fused_adj<{m}, false>"}[kind]


def progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def under_profiler() -> bool:
    return "rocprof" in os.environ.get("LD_PRELOAD", "") or any(k.startswith("ROCPROF") for k in os.environ)


def counter_bytes(rows, counter: str, pattern: str):
    """Mean bytes per dispatch of the kernels whose name contains `pattern`, from the rows of a rocprofv3
    counter_collection.csv holding `counter` (values in KiB per dispatch), scaled by the calibration copy: the first
    `copyBuffer` blit that reads at least 1/4 GiB by the counter is the 1 GiB device-to-device copy, so the scale is
    1 GiB / its value. Returns (bytes, scale, dispatches); ValueError names what is missing."""
    rows = sorted((r for r in rows if r.get("Counter_Name") == counter), key=lambda r: int(r["Dispatch_Id"]))
    if not rows:
        raise ValueError(f"no {counter} rows")
    calib = next((float(r["Counter_Value"]) * 1024 for r in rows
                  if "copyBuffer" in r["Kernel_Name"] and float(r["Counter_Value"]) * 1024 >= (1 << 28)), None)
    if calib is None:
        raise ValueError("no calibration copy")
    vals = [float(r["Counter_Value"]) * 1024 for r in rows if pattern in r["Kernel_Name"]]
    if not vals:
        raise ValueError(f"no dispatch of {pattern}")
    scale = (1 << 30) / calib
    return sum(vals) / len(vals) * scale, scale, len(vals)


def live_traffic(root: str, args, zblock: int, strip: int, kind: str, timeout: float = 150.0):
    """HBM bytes per launch of the `kind` fused instantiation, measured by two child rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE; one counter block each, within what one pass can collect) over this script at
    nt = 8 with the same grid, snapshot mode and block schedule. Each counter is scaled by what it reads for the 1 GiB
    device-to-device copy the child runs first (--copy-calib). Returns (bytes, description) or (None, reason)."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                         else None)
    if prof is None:
        return None, "rocprofv3 not found"
    pat = kernel_pattern(kind, args.snap)
    child = [sys.executable, os.path.join(root, "bench.py"), "--steps", "1", "--warmup", "0", "--n", str(args.n),
             "--nt", "8", "--k", "4", "--snap", args.snap, "--zblock", str(zblock), "--strip", str(strip),
             "--no-plain", "--no-cpu", "--no-batch", "--no-traffic", "--copy-calib"]
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix=f"tti_{counter.lower()}_", dir="/tmp")
        env = dict(os.environ, TMPDIR="/tmp")
        cmd = [prof, "--pmc", counter, "-f", "csv", "-d", out, "-o", "run", "--"] + child
        proc = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, env=env, cwd="/tmp",
                                start_new_session=True)
        try:
            _, err = proc.communicate(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(proc.pid, signal.SIGKILL)
            proc.communicate()
            shutil.rmtree(out, ignore_errors=True)
            return None, f"rocprofv3 {counter} pass exceeded {timeout:.0f} s"
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        rows = []
        for f in files:
            with open(f, newline="") as fh:
                rows += list(csv.DictReader(fh))
        shutil.rmtree(out, ignore_errors=True)
        try:
            per[counter] = counter_bytes(rows, counter, pat)
        except ValueError as e:
            tail = err.decode(errors="replace").strip().splitlines()[-1:] if err else []
            return None, f"rocprofv3 {counter} pass (exit {proc.returncode}): {e}{'; ' + tail[0] if tail else ''}"
    f, w = per["FETCH_SIZE"], per["WRITE_SIZE"]
    return f[0] + w[0], (f"measured in this run: rocprofv3 --pmc FETCH_SIZE x{f[1]:.2f} ({f[2]} launches) + "
                         f"WRITE_SIZE x{w[1]:.2f} ({w[2]} launches), scaled by a 1 GiB copy")


def batched(torch, dev, args):
    """Many shots in one launch per step: `args.batch_shots` seeded uneven apertures of 128-320 cells per axis cut
    from a 1024 x 1024 x 400 model (same distributions as the 512^3 workload, seed 1004), FP16 snapshots every 4
    steps, FP64 images, the z depth of the blocks picked by the autotuner, passes replayed from hipGraphs. Timed:
    forward + adjoint passes of all shots (`value`, interior updates/s), against the same shots run one after the
    other through the same kernels (one launch per shot and step)."""
    import ctypes as C

    import numpy as np

    from miniappsforhpe_amd import autotune
    from miniappsforhpe_amd._lib import call
    from miniappsforhpe_amd.batch import BatchFWI, seeded_apertures
    from miniappsforhpe_amd.grid import Grid
    from miniappsforhpe_amd.models import cfl_dt, layered_tti, ricker
    from miniappsforhpe_amd.propagator import SNAP_FP16

    nt, k = args.batch_nt, 4
    MG = Grid(1024, 1024, 400)
    models, vmax = layered_tti(MG, seed=1004, device=dev)
    dt = cfl_dt(vmax, 10.0, 10.0, 10.0)
    aps = seeded_apertures(MG, args.batch_shots, seed=1004)
    B = BatchFWI(models, MG, aps, ricker(nt, dt, 12.0), 10.0, 10.0, 10.0, dt, nt, k, snap_mode=SNAP_FP16,
                 zblock=128, device=dev)
    del models
    torch.cuda.empty_cache()
    tuned = autotune.tune_batch(B)
    pts = B.interior_points()
    updates = float(pts) * 2 * nt

    def passes():
        B.forward(True)
        B.adjoint(True)

    passes()  # captures the two graphs
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        passes()
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / reps
    # the same shots one at a time (one launch per shot and step), z depth 32 for better fill of small volumes
    descs = B._descs
    handles = []
    for i in range(len(descs)):
        descs[i].zblock = 32
        h = C.c_void_p()
        call("tti_batch_create", C.byref(descs[i]), 1, B.dt, B.dx, B.dy, B.dz, C.byref(h))
        handles.append(h)
    st = torch.cuda.current_stream().cuda_stream

    def sequential():
        for h in handles:
            call("tti_batch_forward", h, nt, k, 0, st)
            call("tti_batch_adjoint", h, nt, k, 0, st)

    sequential()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sequential()
    torch.cuda.synchronize()
    ts = time.perf_counter() - t0
    for h in handles:
        call("tti_batch_destroy", h)
    # per-shot FP64 images summed into the model-sized gradient (once per FWI iteration)
    total = torch.zeros((MG.nz, MG.ny, MG.nx), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sh in B.shots:
        a = sh.ap
        total[a.z0: a.z0 + a.nz, a.y0: a.y0 + a.ny, a.x0: a.x0 + a.nx] += sh.grad
    torch.cuda.synchronize()
    tsum = time.perf_counter() - t0
    sizes = [(a.nx, a.ny, a.nz) for a in aps]
    out = {
        "value": updates / tb, "unit": "grid-pt updates/s", "ms_per_pass_pair": tb * 1e3,
        "sequential_value": updates / ts, "batch_speedup": ts / tb,
        "image_sum_ms": tsum * 1e3, "zblock": tuned["zblock"], "strip": tuned["strip"], "autotune_ms": tuned["ms"],
        "tiles_per_forward_launch": B.tiles,
        "workload": f"{len(aps)} shots, apertures 128-320 per axis (seed 1004) from a 1024x1024x400 model, "
                    f"{nt}+{nt} steps, fp16 snapshots every {k}, FP64 images, 64 receivers per shot",
        "interior_points": pts, "smallest": min(sizes, key=lambda t: t[0] * t[1] * t[2]),
        "largest": max(sizes, key=lambda t: t[0] * t[1] * t[2]),
    }
    B.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); sharding this path over several GPUs "
              "(z slabs with a halo exchange per step, or shots spread over GPUs) is not launched here.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    root = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, root)
    from miniappsforhpe_amd.grid import Grid
    from miniappsforhpe_amd.models import cfl_dt, layered_tti, ricker
    from miniappsforhpe_amd.propagator import SNAP_BF16, SNAP_FP16, SNAP_FP32, SnapshotStore, TTIPropagator
    from miniappsforhpe_amd import ops

    dev = torch.device("cuda:0")
    if args.copy_calib:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        c = torch.empty_like(a)
        c.copy_(a)
        torch.cuda.synchronize()
        del a, c
    n, nt, k = args.n, args.nt, args.k
    G = Grid(n, n, n)
    dx = dy = dz = 10.0
    models, vmax = layered_tti(G, seed=1002, device=dev)
    dt = cfl_dt(vmax, dx, dy, dz)
    P = TTIPropagator(G, models, dx, dy, dz, dt, device=dev, zblock=args.zblock or 128, strip=args.strip)
    del models
    tuned = None
    if args.zblock == 0:
        from miniappsforhpe_amd import autotune

        tuned = autotune.tune_propagator(P)
        progress(f"schedule: zblock {P.zblock}, strip {P.strip}")
    src = (G.ghost + n // 2) + G.ldx * ((G.ghost + n // 2) + G.ldy * (G.ghost + n // 2))
    wav = ricker(nt, dt, 12.0).tolist()
    nsnap = (nt + k - 1) // k
    mode = {"fp16": SNAP_FP16, "bf16": SNAP_BF16, "fp32": SNAP_FP32}[args.snap]
    store = SnapshotStore(G, mode, nsnap, dev)
    grad = torch.zeros((n, n, n), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    # launch timings per kernel instantiation (HIP events on the launch stream)
    events = {"fwd": [], "fwd+snapshot": [], "adj": [], "adj+imaging": []}

    def fused_pass(record=False):
        P.reset()
        P.prime()
        for it in range(nt):
            snap = store.desc(it // k) if it % k == 0 else None
            if record:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
            P.step_fwd_fused(src, dt * wav[it], snap)
            if record:
                e1.record(stream)
                events["fwd+snapshot" if snap is not None else "fwd"].append((e0, e1))
        for it in reversed(range(nt)):
            snap = store.desc(it // k) if it % k == 0 else None
            if record:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
            P.step_adj_fused(snap, grad if snap is not None else None)
            if record:
                e1.record(stream)
                events["adj+imaging" if snap is not None else "adj"].append((e0, e1))

    updates = float(n) ** 3 * 2 * nt
    for _ in range(args.warmup):
        fused_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        fused_pass(record=(s == args.steps - 1))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = updates * args.steps / (t1 - t0)
    progress(f"fused: {args.steps} steps, {ms_per_step:.1f} ms/step")

    mean_ms = {key: sum(a.elapsed_time(b) for a, b in ev) / len(ev) for key, ev in events.items() if ev}
    total_ms = {key: sum(a.elapsed_time(b) for a, b in ev) for key, ev in events.items() if ev}
    dominant = max(total_ms, key=total_ms.get)  # the instantiation that takes most of the pass
    kernel_names = {
        "fwd": "k_fused_fwd<4, TTI_SNAP_NONE, false> (tti_fused_fwd_step)",
        "fwd+snapshot": f"k_fused_fwd<4, {args.snap} snapshot, false> (tti_fused_fwd_step)",
        "adj": "k_fused_adj<TTI_SNAP_NONE, false> (tti_fused_adj_step)",
        "adj+imaging": f"k_fused_adj<{args.snap} imaging, false> (tti_fused_adj_step)",
    }
    dom_ms = mean_ms[dominant]
    alg_bytes = BYTES_PER_UPDATE * float(n) ** 3
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    # HBM traffic of the dominant kernel per launch: measured live by rocprofv3 counter passes in child processes;
    # else the committed counter summary of this workload
    traffic, traffic_src = None, None
    live_err = "skipped (--no-traffic or run under a profiler)"
    if not args.no_traffic and not under_profiler():
        traffic, live_err = live_traffic(root, args, P.zblock, P.strip, dominant)
        progress(f"traffic of {dominant}: {traffic} ({live_err})")
        if traffic is not None:
            traffic_src, live_err = live_err, None
    tfile = os.path.join(root, "profiles", f"fused_traffic_{n}.json")
    if traffic is None and os.path.exists(tfile):
        try:
            with open(tfile) as f:
                tr = json.load(f)
            if dominant in tr.get("kernels", {}):
                traffic = tr["kernels"][dominant]["traffic_bytes"]
                traffic_src = (f"profiles/fused_traffic_{n}.json (rocprofv3 FETCH_SIZE x{tr['fetch_scale']:.2f} + "
                               f"WRITE_SIZE x{tr['write_scale']:.2f}, scaled by a plain copy of known size); live "
                               f"measurement: {live_err}")
        except (OSError, ValueError, KeyError):
            traffic, traffic_src = None, None

    # the box's measured streaming bandwidth: a 2 GiB device-to-device copy (read + write bytes / time), the
    # practical ceiling next to the 8 TB/s specification
    src_buf = torch.empty(1 << 29, dtype=torch.float32, device=dev)
    dst_buf = torch.empty_like(src_buf)
    dst_buf.copy_(src_buf)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(5):
        dst_buf.copy_(src_buf)
    c1.record()
    c1.synchronize()
    stream_gbs = 5 * 2 * src_buf.numel() * 4 / (c0.elapsed_time(c1) * 1e-3) / 1e9
    del src_buf, dst_buf
    torch.cuda.empty_cache()

    result = {
        "metric": "grid-pt updates/s, TTI 8th-order fwd+adj 512^3",
        "value": value,
        "unit": "grid-pt updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"TTI fwd+adj+zero-lag imaging, {n}^3 FP32 + 4-pt ghosts, {nt}+{nt} steps/step, "
                        f"{args.snap} snapshots every {k} steps, FP64 image, layered vp 1500-4500 + smoothed TTI "
                        f"params (seed 1002), 12 Hz Ricker",
            "grid": [n, n, n], "nt": nt, "xcorr_step": k, "snapshot": args.snap, "dt": dt, "dx": dx,
            "parallelism": "single GPU", "zblock": P.zblock, "strip": P.strip,
            "autotune_ms": tuned["ms"] if tuned else None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_names[dominant],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "stream_copy_GBs": stream_gbs,
            "frac_of_stream_copy": achieved / stream_gbs,
            "traffic": traffic,
            "traffic_source": traffic_src,
            # the bytes the kernel moves per second, against the copy bandwidth of this box
            "traffic_GBs": traffic / (dom_ms * 1e-3) / 1e9 if traffic else None,
            "traffic_frac_of_stream_copy": traffic / (dom_ms * 1e-3) / 1e9 / stream_gbs if traffic else None,
            "algorithmic_bytes_per_launch": alg_bytes,
            "mean_ms": mean_ms,
        },
        "snapshot_store_bytes": store.nbytes(),
    }

    if not args.no_plain:
        del store
        fstore = SnapshotStore(G, SNAP_FP32, nsnap, dev)
        g32 = torch.zeros((n, n, n), dtype=torch.float32, device=dev)

        def plain_pass():
            P.reset()
            for it in range(nt):
                P.step_fwd_plain(src, dt * wav[it])
                if it % k == 0:
                    fstore.u[it // k].copy_(G.interior(P.u))
                    fstore.w[it // k].copy_(G.interior(P.w))
            for it in reversed(range(nt)):
                P.step_adj_plain()
                if it % k == 0:
                    ops.vp_gradient(G, fstore.u[it // k], fstore.w[it // k], g32, P.u, P.w)

        if G.nx % 64 == 0:
            plain_pass()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            plain_pass()
            torch.cuda.synchronize()
            tp = time.perf_counter() - t0
            pv = updates / tp
            result["plain_fp32"] = {"value": pv, "ms_per_step": tp * 1e3, "fused_speedup": value / pv}
            progress(f"plain: {tp * 1e3:.1f} ms/step")

    if not args.no_batch:
        P = store = fstore = g32 = grad = None  # release the 512^3 buffers
        torch.cuda.empty_cache()
        result["batched"] = batched(torch, dev, args)
        progress("batched done")

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(root, "oracle"))
        try:
            import oracle as O

            ncpu = 128
            sec, th = O.bench_fwdadj(ncpu, 10)  # calibration: seconds per forward + adjoint step
            steps = max(2, int(10 * args.cpu_seconds / max(sec, 1e-3)))
            sec, th = O.bench_fwdadj(ncpu, steps)
            result["cpu_baseline"] = {
                "value": ncpu**3 * 2 * steps / sec,
                "unit": "grid-pt updates/s",
                "cores": th,
                "kind": "port",
                "sample": f"{steps} forward + {steps} adjoint steps of the FP32 OpenMP CPU restatement "
                          f"(reference step structure) on a {ncpu}^3 interior, {sec:.1f} s",
            }
        except Exception as e:  # the baseline is reported, never required
            result["cpu_baseline"] = {"value": None, "error": str(e)}
        progress("cpu baseline done")
        result["etti_reference_cpu"] = etti_reference_cpu(root)

    print(json.dumps(result))


def etti_reference_cpu(root: str) -> dict:
    """The reference's own CPU code beside its GPU twin: the reference's CPU driver of the staggered elastic TTI
    system (cpuMiniApp/simulator4, built from the reference's archive into oracle/_ref) on 8 MPI ranks (2 x 2 x 2,
    one per core) against this library's simulator3SetUp on the GPU, both on 3 x 3 x 3 zones with npoints 64 and
    npml 16 (96^3 points), 24 forward + 24 adjoint steps, snapshots every 4. Rates are grid-point updates/s of each
    driver's execute(), its own timing ("Total Time" of the reference's output; wall time around execute for the
    GPU driver). The reference's CPU code has no counterpart of the TTI pseudo-acoustic path, so this is a side
    measurement, not the cpu_baseline."""
    import re
    import subprocess
    import time

    exe = os.path.join(root, "oracle", "_ref", "miniapp_simulator4")
    mpiexec = "/opt/conda/bin/mpiexec"
    npts, npml, niter, k, ranks = 64, 16, 24, 4, 8
    upd = float(npts + 2 * npml) ** 3 * 2 * niter
    out = {"config": f"3x3x3 zones, npoints {npts}, npml {npml} ({npts + 2 * npml}^3 points), {niter}+{niter} steps, "
                     f"snapshots every {k}"}
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        out["error"] = "oracle/_ref/miniapp_simulator4 or mpiexec missing"
        return out
    try:
        p = subprocess.run([mpiexec, "-n", str(ranks), exe, "2", "2", "2", str(npts), str(npml), str(niter), str(k)],
                           capture_output=True, text=True, timeout=300, cwd="/tmp")
        m = re.search(r"Total Time:\s*([0-9.eE+-]+)", p.stdout)
        if p.returncode != 0 or not m:
            out["error"] = f"reference CPU driver failed ({p.returncode}): {p.stdout[-300:]} {p.stderr[-300:]}"
            return out
        cpu_s = float(m.group(1))
        out["cpu"] = {"value": upd / cpu_s, "unit": "grid-pt updates/s", "cores": ranks, "kind": "reference",
                      "seconds": cpu_s, "driver": "cpuMiniApp/simulator4 (reference, g++ -O3, mpich), 8 MPI ranks"}
        import torch

        from miniappsforhpe_amd.host import Simulator3

        with Simulator3(npts, npml, niter, k, (3, 3, 3)) as sim:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sim.execute()
            torch.cuda.synchronize()
            gpu_s = time.perf_counter() - t0
        out["gpu"] = {"value": upd / gpu_s, "unit": "grid-pt updates/s", "seconds": gpu_s,
                      "driver": "simulator3SetUp::execute on this library (one MI355X, host snapshots)"}
        out["gpu_over_cpu"] = cpu_s / gpu_s
    except Exception as e:  # reported, never required
        out["error"] = str(e)
    return out


if __name__ == "__main__":
    main()
