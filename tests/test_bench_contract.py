"""bench.py's contract pieces that run without a GPU: the one-GPU refusal
and the figures it takes from the committed rocprofv3 summaries of the
single-run workloads (HBM traffic per launch, counter-based FP32 work)."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _counter(name, workload="c2_hex"):
    with open(os.path.join(ROOT, "profiles", bench.PROFILES[workload][0])) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2 and parts[0] == name:
                try:
                    return float(parts[1].replace(",", ""))
                except ValueError:
                    continue
    raise AssertionError(name)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]


def test_traffic_is_the_calibrated_counters_of_the_summary():
    for wl, (_, walkers, steps) in bench.PROFILES.items():
        t = bench._profiled_traffic(bench._profile(wl, walkers, steps))
        assert t == int(_counter("FETCH_SIZE", wl) * 1024 * 2 + _counter("WRITE_SIZE", wl) * 1024), wl
    assert bench._profile("c2_hex", 1000, 250000) is None      # other sizes: not measured
    assert bench._profiled_traffic(None) is None


def test_counted_fp32_work_is_64_lanes_of_the_flop_counter():
    for wl, (_, walkers, steps) in bench.PROFILES.items():
        path = bench._profile(wl, walkers, steps)
        flop = bench._profiled_valu_flops(path)
        assert flop == 64.0 * _counter("SQ_INSTS_VALU_FLOPS_FP32", wl)
        c = bench._counted(flop, 1000.0, path)
        assert abs(c["tflops"] - flop / 1e12) < 1e-9 and abs(c["frac"] - c["tflops"] / bench.PEAK_FP32_TFLOPS) < 1e-12
    assert bench._counted(None, 1000.0, None) is None
