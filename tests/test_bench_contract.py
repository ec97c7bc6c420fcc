"""bench.py's host logic on the CPU: the one-GPU guard and the rocprofv3 counter arithmetic of the traffic figure."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _write(path, rows):
    cols = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for i, (name, counter, kib) in enumerate(rows):
            w.writerow({"Correlation_Id": i + 1, "Dispatch_Id": i + 1, "Kernel_Name": name, "Counter_Name": counter,
                        "Counter_Value": kib})


def test_counter_bytes_scales_by_the_first_large_copy(tmp_path):
    adj = "void tti::fused::k_fused_adj<0, false>(tti::fused::Box, ...)"
    img = "void tti::fused::k_fused_adj<2, false>(tti::fused::Box, ...)"
    p = tmp_path / "run_counter_collection.csv"
    # a small blit (ignored), the 1 GiB calibration copy read as 512 MiB, two adjoint launches, one imaging launch,
    # then a later 2 GiB copy that must not be taken as the calibration
    _write(p, [("__amd_rocclr_copyBuffer", "FETCH_SIZE", 64.0),
               ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 512.0 * 1024),
               (adj, "FETCH_SIZE", 8.0e6), (adj, "FETCH_SIZE", 9.0e6), (img, "FETCH_SIZE", 1.0e7),
               ("__amd_rocclr_copyBuffer", "FETCH_SIZE", 2048.0 * 1024)])
    b, scale, n = bench.counter_bytes(_rows(p), "FETCH_SIZE", bench.kernel_pattern("adj", "fp16"))
    assert scale == 2.0 and n == 2 and b == pytest.approx(8.5e6 * 1024 * 2.0)
    b, _, n = bench.counter_bytes(_rows(p), "FETCH_SIZE", bench.kernel_pattern("adj+imaging", "fp16"))
    assert n == 1 and b == pytest.approx(1.0e7 * 1024 * 2.0)
    with pytest.raises(ValueError, match="no WRITE_SIZE rows"):
        bench.counter_bytes(_rows(p), "WRITE_SIZE", "k_fused_adj")
    with pytest.raises(ValueError, match="no dispatch"):
        bench.counter_bytes(_rows(p), "FETCH_SIZE", bench.kernel_pattern("fwd", "fp16"))
    q = tmp_path / "nocal.csv"
    _write(q, [(adj, "FETCH_SIZE", 8.0e6)])
    with pytest.raises(ValueError, match="calibration"):
        bench.counter_bytes(_rows(q), "FETCH_SIZE", "k_fused_adj")


def test_kernel_patterns_name_the_timed_instantiations():
    # forward <block rows, snapshot mode, sponge, sigma in the snapshot>; adjoint <snapshot mode, sponge>; FP16 is mode 2
    assert bench.kernel_pattern("fwd", "fp16") == "k_fused_fwd<4, 0, false,"
    assert bench.kernel_pattern("fwd+snapshot", "fp16") == "k_fused_fwd<4, 2, false,"
    assert bench.kernel_pattern("adj+imaging", "bf16") == "k_fused_adj<3, false>"


def test_live_traffic_without_rocprofv3_reports_why(monkeypatch):
    import shutil
    import types

    monkeypatch.setattr(shutil, "which", lambda name: None)
    real_exists = os.path.exists
    monkeypatch.setattr(os.path, "exists", lambda p: False if "rocprofv3" in str(p) else real_exists(p))
    args = types.SimpleNamespace(n=64, snap="fp16")
    traffic, why = bench.live_traffic(ROOT, args, 64, 0, "adj")
    assert traffic is None and "not found" in why


def test_under_profiler_detects_the_preload(monkeypatch):
    monkeypatch.setenv("LD_PRELOAD", "/opt/rocm/lib/librocprofiler-sdk-tool.so")
    assert bench.under_profiler()
    monkeypatch.delenv("LD_PRELOAD")
    for k in [k for k in os.environ if k.startswith("ROCPROF")]:
        monkeypatch.delenv(k)
    assert not bench.under_profiler()
