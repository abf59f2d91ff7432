"""bench.py's host-side pieces, on the CPU: the command-line refusal of several GPUs and the
per-workload CPU baseline (the reference library built from its sources, when present)."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "c-permutation-avoiding-convolution_amd"))


def test_several_gpus_are_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU only" in r.stdout + r.stderr


def test_cpu_baseline_workloads_report_every_workload():
    import bench
    import synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import RefLib

    if not RefLib.available():
        pytest.skip("oracle/_ref not built")
    gpu = {"1d_65536": 0.02, "2d_2048x2048": 0.15, "2d_4096x256": None, "batched_4096": 0.2}
    r = bench.cpu_baseline_workloads(synth, gpu)
    assert r["kind"] == "reference" and r["cores"] == 1
    w = r["workloads"]
    assert set(w) == {"1d_65536", "2d_2048x2048", "2d_4096x256", "batched_4096"}
    assert w["1d_65536"]["radix"] == 2 and w["2d_4096x256"]["radix"] == 4
    for name, e in w.items():
        sec = e.get("s_per_conv", e.get("s_per_batch"))
        assert sec > 0.0
        if gpu[name] is None:
            assert e["gpu_speedup"] is None
        else:
            assert e["gpu_speedup"] == pytest.approx(sec * 1e3 / gpu[name])


def test_cpu_baseline_runs_pinned_to_one_core():
    """The CPU baseline runs in a child process pinned to one core (the reference's taskset -c 0) and
    reports the host's CPU model, core counts and the reference library's compile flags."""
    import bench
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import RefLib

    if not RefLib.available():
        pytest.skip("oracle/_ref not built")
    r = bench.cpu_baseline(1)
    assert len(r["pinned_cores"]) == 1 and r["cores"] == 1
    assert r["warmup_calls"] == 1 and r["timed_calls"] == 1 and r["value"] > 0
    assert r["kind"] == "reference" and "-O3" in r["compile_flags"] and "-ffast-math" in r["compile_flags"]
    assert r["host_cores"] >= 1 and r["cpu_model"]


def test_traffic_figure_names_its_source():
    import bench

    t, src = bench.measured_traffic("f32_fused_s_4x8x8_h_l32")
    assert t and t > 0 and "profiles/traffic_3d256_fp32.json" in src
    t, src = bench.measured_traffic("no_such_kernel")
    assert t is None and "no entry" in src
