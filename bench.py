#!/usr/bin/env python3
"""Benchmark: walker-steps/s of the gfx950 random-walk kernel on the
hexagonal-lattice cosine-OGSE workload (1000 permeable cylinders, 1e6
walkers, 250,000 steps; see axonexa_amd/workloads.py::c2_hex), with the
CPU oracle timed beside it on the host cores.

One bench step = one full walk (init + every time step of every walker)
plus the fp64 signal reduction, inputs resident on the device.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

--workload c3_gamma measures the gamma-packing workload (400 cylinders,
4e6 walkers x 80,000 steps, 64 profiles x 3 axes); c1_free / c1_cylinder the
small PGSE workloads; c4_batch the batched uneven workload; c5_surrogate the
learned surrogate (data, training, held-out error, speed); mag_lf /
mag_stream the magnetization path.

Prints one JSON line.  The default run (hexagonal workload) adds a
"secondary" object after the primary measurement: one or two timed runs of
each other workload in the same process (about 30 s; --no-secondary skips
it).
"""
import argparse
import contextlib
import dataclasses
import io
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3     # MI355X vector FP32 (spec sheet)
PEAK_HBM_GBS = 8000.0


def _cpu_counts():
    """(threads used, nproc, hardware_concurrency): the CPU baseline runs one
    thread per processing unit `nproc` grants this process (it honours the
    affinity mask and OMP_NUM_THREADS, the box's CPU share);
    hardware_concurrency = the online CPUs (what std::thread reports)."""
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.split()[0])
    except (OSError, ValueError, IndexError, subprocess.CalledProcessError):
        nproc = len(os.sched_getaffinity(0))
    hw = os.cpu_count() or 1
    return max(1, nproc), nproc, hw


def _cpu_threads() -> int:
    return _cpu_counts()[0]


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# rocprofv3 summaries of the walk kernel of the single-run workloads at their
# full size: (file, walkers, time steps)
PROFILES = {
    "c2_hex": ("walk_kernel_c2_hex.txt", 1000000, 250000),
    "c3_gamma": ("walk_kernel_c3_gamma.txt", 4000000, 80000),
}


def _profile(workload: str, walkers: int, n_steps: int):
    """Path of the workload's profile summary when this run has its size."""
    if workload not in PROFILES:
        return None
    name, w, n = PROFILES[workload]
    path = os.path.join(ROOT, "profiles", name)
    return path if (walkers, n_steps) == (w, n) and os.path.exists(path) else None


def _counted(flop, kernel_ms, path):
    """Counter-based FP32 throughput of this run's walk kernel."""
    if flop is None:
        return None
    tf = flop / (kernel_ms * 1e-3) / 1e12
    return {"flop_per_launch": flop, "tflops": tf, "frac": tf / PEAK_FP32_TFLOPS,
            "source": f"64 x SQ_INSTS_VALU_FLOPS_FP32 per launch from profiles/{os.path.basename(path)}, "
                      "over this run's kernel time"}


def _profiled_traffic(path):
    """HBM bytes per walk launch measured with rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (calibrated against a plain copy) for the workload at its
    full size; None when there is no summary for this run."""
    if path is None:
        return None
    with open(path) as f:
        for line in f:
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    return None


def _profiled_valu_flops(path):
    """FP32 VALU work per walk launch as the counters count it: 64 lanes x
    SQ_INSTS_VALU_FLOPS_FP32 (ADD + MUL + 2 FMA + transcendental wave
    instructions) from the workload's rocprofv3 summary; None when there is
    no summary for this run."""
    if path is None:
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2 and parts[0] == "SQ_INSTS_VALU_FLOPS_FP32":
                try:
                    return 64.0 * float(parts[1])
                except ValueError:   # the counter's name in a command line
                    continue
    return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2_hex",
                    choices=["c2_hex", "c3_gamma", "c1_free", "c1_cylinder", "c4_batch",
                             "c5_surrogate", "mag_lf", "mag_stream"])
    ap.add_argument("--mag-acqs", type=int, default=1000, help="mag_stream: acquisitions")
    ap.add_argument("--lds-budget", type=int, default=0,
                    help="c4_batch: LDS bytes per workgroup for substrate tables (0: library default)")
    ap.add_argument("--mag-fields", type=int, default=2, help="mag_lf: fields per walker")
    ap.add_argument("--mag-integrator", type=int, default=0, help="mag_lf: 0 rotate, 1 RK4, 2 DOPRI")
    ap.add_argument("--sur-geoms", type=int, default=8448,
                    help="c5_surrogate: training geometries (8 queries of 30 signals each: "
                         "8448 -> 2.03e6 signals)")
    ap.add_argument("--sur-steps", type=int, default=4000, help="c5_surrogate: Adam steps")
    ap.add_argument("--no-unbatched", action="store_true",
                    help="c4_batch: skip timing the instances one plan at a time")
    ap.add_argument("--scale", type=float, default=1.0, help="c4_batch: walker-count scale")
    ap.add_argument("--walkers", type=int, default=0, help="0: the workload's own count")
    ap.add_argument("--n-steps", type=int, default=0, help="time steps per walk (0: the workload's)")
    ap.add_argument("--cpu-walkers", type=int, default=32768,
                    help="walkers in the CPU-baseline sample (full step count)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0, help="threads per block (0: plan default)")
    ap.add_argument("--task-steps", type=int, default=0,
                    help="steps per walk task (0: automatic, -1: the whole walk per task)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="c2_hex: skip the short measurements of the other workloads")
    args = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(walkers are independent: a multi-GPU run would give each "
                                   "device its own walker-id range)"}))
        return 2

    if args.workload == "c4_batch":
        return bench_batch(args)
    if args.workload == "c5_surrogate":
        return bench_surrogate(args)
    if args.workload == "mag_lf":
        return bench_mag(args)
    if args.workload == "mag_stream":
        return bench_mag_stream(args)

    out = bench_single(args)
    if args.workload == "c2_hex" and not args.no_secondary:
        out["secondary"] = _secondary(args)
    print(json.dumps(out))
    return 0


def _quiet(fn, args) -> dict:
    """Run one of the bench_* functions that print a JSON line and return
    that line as a dict instead."""
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        rc = fn(args)
    lines = [l for l in buf.getvalue().splitlines() if l.startswith("{")]
    if rc or not lines:
        raise RuntimeError(f"{fn.__name__} returned {rc}")
    return json.loads(lines[-1])


def _surrogate_forward(batch: int = 65536) -> dict:
    """The surrogate network's forward pass at `batch` queries with random
    weights (its speed does not depend on the weights), median of 3 timed
    passes with HIP events."""
    import numpy as np

    from axonexa_amd import engine

    dims = [268] + [1024] * 6 + [30]
    rng = np.random.default_rng(0)
    with engine.MLP(dims, max_batch=batch) as m:
        m.set_params(np.concatenate([np.concatenate([rng.uniform(-1, 1, a * b) / np.sqrt(a), np.zeros(b)])
                                     for a, b in zip(dims[:-1], dims[1:])]).astype(np.float32))
        m.forward(rng.uniform(-1, 1, size=(batch, 268)).astype(np.float32))
        ms = float(np.median([m.time_forward(batch) for _ in range(3)]))
    flops = 2.0 * batch * sum(a * b for a, b in zip(dims[:-1], dims[1:]))
    tf = flops / (ms * 1e-3) / 1e12
    return {"value": batch / (ms * 1e-3), "unit": "queries/s", "ms_per_step": ms,
            "roofline": {"bound": "mfma", "achieved": tf, "peak": 2500.0, "unit": "TFLOP/s", "frac": tf / 2500.0},
            "config": f"MLP 268 -> 1024 x 6 -> 30, BF16, batch {batch}, forward only, random weights "
                      "(training and held-out error: bench.py --workload c5_surrogate)"}


def _secondary(args) -> dict:
    """Short measurements of the other workloads in the same process, after
    the primary figure: one or two timed runs each, no CPU leg.  Each entry
    keeps the workload's own metric; the full runs with their CPU baselines
    are `bench.py --workload <name>`."""
    keep = ("value", "unit", "ms_per_step", "steps", "warmup")
    out = {}

    def entry(d):
        e = {k: d[k] for k in keep if k in d}
        r = d.get("roofline") or {}
        e["roofline_frac"] = r.get("frac")
        e["roofline_bound"] = r.get("bound")
        e["config"] = (d.get("config") or {}).get("workload")
        return e

    def ns(**kw):
        d = dict(vars(args))
        d.update(no_cpu=True, no_unbatched=True, walkers=0, n_steps=0, block=0, no_secondary=True)
        d.update(kw)
        return argparse.Namespace(**d)

    runs = [
        ("c3_gamma", bench_single, ns(workload="c3_gamma", steps=1, warmup=0)),
        ("c1_free", bench_single, ns(workload="c1_free", steps=2, warmup=1)),
        ("c1_cylinder", bench_single, ns(workload="c1_cylinder", steps=2, warmup=1)),
        ("c4_batch", bench_batch, ns(workload="c4_batch", steps=1, warmup=0)),
        ("mag_lf", bench_mag, ns(workload="mag_lf", steps=1, warmup=1)),
    ]
    for name, fn, a in runs:
        t0 = time.perf_counter()
        d = fn(a) if fn is bench_single else _quiet(fn, a)
        out[name] = dict(entry(d), seconds=time.perf_counter() - t0)
    t0 = time.perf_counter()
    out["c5_surrogate_forward"] = dict(_surrogate_forward(), seconds=time.perf_counter() - t0)
    return out


def bench_single(args) -> dict:
    """One plan of a single-substrate workload (hexagonal lattice, gamma
    packing, free box, single cylinder): warmup walks, then `steps` timed
    walks + reductions, with the CPU oracle beside it."""
    import numpy as np

    from axonexa_amd import abi, engine
    from axonexa_amd import workloads as W

    make = {"c2_hex": W.c2_hex, "c3_gamma": W.c3_gamma, "c1_free": W.c1_free,
            "c1_cylinder": W.c1_cylinder}[args.workload]
    kw = {}
    if args.walkers:
        kw["n_walkers"] = args.walkers
    if args.n_steps:
        kw["n_steps"] = args.n_steps
    wl = make(**kw)
    n = wl.acq.n_steps
    plan = engine.Plan(wl.sub, wl.acq, wl.run, block=args.block, lds_budget=args.lds_budget)
    plan.set_schedule(args.task_steps)
    cfg = plan.config()

    for _ in range(args.warmup):
        plan.walk(0, n).reduce().sync()

    kernel_ms = []
    plan.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(plan.time_walk(0, n))
        plan.reduce()
        plan.sync()
    t1 = time.perf_counter()
    sig = plan.results()
    elapsed = t1 - t0
    ws = wl.walker_steps
    value = ws * args.steps / elapsed
    avg_ms = float(np.mean(kernel_ms))
    rho = sig.collisions / ws
    F = wl.flops_per_walker_step(rho)
    achieved = F * ws / (avg_ms * 1e-3) / 1e12
    prof = _profile(args.workload, wl.run.n_walkers, n)

    out = {
        "metric": "walker-steps/s",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": wl.name + ": " + wl.description,
            "walkers": wl.run.n_walkers,
            "time_steps": n,
            "dt_ms": wl.acq.dt,
            "waveforms": wl.acq.n_waveforms,
            "profiles": wl.acq.n_profiles,
            "primitives": wl.sub.n_prims,
            "block": cfg["block"],
            "bin_width_um": round(cfg["bin_width"], 4),
            "lds_bytes": cfg["lds_bytes"],
            "unit_cells": cfg.get("unit_cells", 1),
            "task_steps": plan.schedule()["last_task_steps"],
            "grid": plan.schedule()["last_grid"],
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": _profiled_traffic(prof),
            "traffic_source": (f"rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of this workload "
                               f"(profiles/{os.path.basename(prof)}), not measured in this run" if prof else None),
            "flops_per_walker_step": F,
            "walker_steps_per_launch": ws,
            "kernel_ms": avg_ms,
            "valu_fp32_counted": _counted(_profiled_valu_flops(prof), avg_ms, prof),
            "collisions_per_walker_step": rho,
        },
        "signal": {
            "E_first": float(sig.E[0]),
            "E_last": float(sig.E[-1]),
            "transmissions": sig.transmissions,
            "cap_hits": sig.cap_hits,
        },
    }

    if not args.no_cpu:
        from oracle import pyoracle as po

        nw = min(args.cpu_walkers, wl.run.n_walkers)
        sample = make(**dict(kw, n_walkers=nw))
        threads = _cpu_threads()
        tc = time.perf_counter()
        ost = po.walk(sample.sub, sample.acq, sample.run, threads=threads,
                      bin_width=cfg["bin_width"])
        tcpu = time.perf_counter() - tc
        cpu_rate = sample.walker_steps / tcpu
        osig = po.reduce(sample.sub, sample.acq, sample.run, ost)
        # the same walkers on the GPU: walker-by-walker parity of the sample
        with engine.Plan(sample.sub, sample.acq, sample.run) as sp:
            sp.walk(0, n).reduce()
            gsig = sp.results()
            gst = sp.get_state()
        bit = np.all(gst.reshape(-1, nw) == ost.reshape(-1, nw), axis=0)
        dE = np.abs(gsig.E_re - osig.E_re) / np.maximum(osig.sigma_re, 1e-300)
        _, nproc, hw = _cpu_counts()
        out["cpu_baseline"] = {
            "value": cpu_rate,
            "unit": "walker-steps/s",
            "cores": threads,
            "nproc": nproc,
            "hardware_concurrency": hw,
            "kind": "port",
            "sample": f"{nw} walkers x {n} steps of the same workload (walker ids 0..{nw - 1}), "
                      f"oracle with the GPU's cell bins, {threads} threads on {_cpu_model()}",
            "seconds": tcpu,
        }
        out["vs_cpu"] = value / cpu_rate
        # the reference-faithful variant: every step scans every primitive
        # image of the whole cell (lattice.cuh:184-199), on a shorter sample
        lw = min(256, wl.run.n_walkers)
        ln = min(n, 20000)
        lin = make(**dict(kw, n_walkers=lw, n_steps=ln))
        tl = time.perf_counter()
        with po.whole_cell():
            po.walk(lin.sub, lin.acq, lin.run, threads=threads, bin_width=0.0)
        tlin = time.perf_counter() - tl
        out["cpu_baseline_linear_scan"] = {
            "value": lin.walker_steps / tlin,
            "unit": "walker-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{lw} walkers x {ln} steps of the same workload, oracle scanning all primitive images "
                      f"of the whole cell every step, {threads} threads",
            "seconds": tlin,
        }
        out["signal_error_vs_oracle"] = {
            "sample_walkers": nw,
            "bitwise_equal_walkers": float(bit.mean()),
            "collisions_equal": bool(gsig.collisions == osig.collisions),
            "transmissions_equal": bool(gsig.transmissions == osig.transmissions),
            "max_abs_dE_over_sigma": float(np.max(dE)),
        }
    plan.close()
    return out


PEAK_FP64_TFLOPS = 78.6      # MI355X vector FP64 (spec sheet)


def bench_mag(args) -> int:
    """Magnetization (Bloch) path: field updates/s of the fp64 rotation walk
    on the low-field saddle-coil workload, with the CPU oracle beside it."""
    import numpy as np

    from axonexa_amd import abi, engine
    from axonexa_amd import workloads as W

    kw = {"n_fields": args.mag_fields, "integrator": args.mag_integrator}
    if args.walkers:
        kw["n_walkers"] = args.walkers
    wl = W.mag_lf(**kw)
    n = wl.field.n_steps
    m = engine.Mag(wl.sub, wl.field, wl.run)
    m.set_schedule(getattr(args, "task_steps", 0))
    for _ in range(args.warmup):
        m.walk(0, n)
        m.results()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(m.time_walk(0, n))
        res = m.results()
    elapsed = time.perf_counter() - t0
    value = wl.field_updates * args.steps / elapsed
    avg_ms = float(np.mean(kernel_ms))
    # the timed walk covers steps 1 .. n-1 (placement and t = 0 are outside)
    upd = wl.run.n_walkers * (n - 1) * wl.field.n_fields
    achieved = W.MAG_ROTATE_FLOPS * upd / (avg_ms * 1e-3) / 1e12 if args.mag_integrator == 0 else None
    out = {
        "metric": "field-updates/s",
        "value": value,
        "unit": "field-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": wl.name + ": " + wl.description, "walkers": wl.run.n_walkers,
                   "time_steps": n, "fields": wl.field.n_fields, "integrator": args.mag_integrator},
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS if achieved else None,
            "traffic": None,
            "flops_per_field_update": W.MAG_ROTATE_FLOPS if args.mag_integrator == 0 else None,
            "field_updates_per_launch": upd,
            "kernel_ms": avg_ms,
            "walker_steps_per_s_kernel": wl.run.n_walkers * (n - 1) / (avg_ms * 1e-3),
        },
        "signal": {"transverse": [float(v) for v in res.transverse], "collisions": res.collisions},
    }
    if not args.no_cpu:
        from oracle import pyoracle as po
        # 28 x 32,768 walkers by default = the whole workload: about 5 s (2 fields)
        # to 20 s (8 fields) of oracle time on 16 host threads
        nw = min(28 * args.cpu_walkers, wl.run.n_walkers)
        sample = W.mag_lf(**dict(kw, n_walkers=nw))
        threads = _cpu_threads()
        tc = time.perf_counter()
        ost = po.walk_mag(sample.sub, sample.field, sample.run, threads=threads)
        tcpu = time.perf_counter() - tc
        with engine.Mag(sample.sub, sample.field, sample.run) as sm:
            sm.walk()
            gst = sm.get_state(0)
        bit = np.all(gst.reshape(-1, nw) == ost.reshape(-1, nw), axis=0)
        out["cpu_baseline"] = {
            "value": sample.field_updates / tcpu,
            "unit": "field-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{nw} walkers x {n} steps x {sample.field.n_fields} fields of the same workload, "
                      f"{threads} threads on {_cpu_model()}",
            "seconds": tcpu,
        }
        out["vs_cpu"] = value / out["cpu_baseline"]["value"]
        out["sample_bitwise_equal_walkers"] = float(bit.mean())
    print(json.dumps(out))
    m.close()
    return 0


def bench_mag_stream(args) -> int:
    """Every acquisition of the low-field experiment in one batched launch
    (axx_mag_batch_*), against the same acquisitions run one by one."""
    import numpy as np

    from axonexa_amd import engine
    from axonexa_amd import workloads as W

    items = W.mag_lf_stream(n_acq=args.mag_acqs)
    upd = sum(r.n_walkers * f.n_steps * f.n_fields for _, f, r in items)
    walker_steps = sum(r.n_walkers * (f.n_steps - 1) for _, f, r in items)
    tc = time.perf_counter()
    b = engine.MagBatch(items)
    t_create = time.perf_counter() - tc
    for _ in range(args.warmup):
        b.run()
        b.results(0)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(b.time())
        last = b.results(len(items) - 1)
    elapsed = time.perf_counter() - t0
    value = upd * args.steps / elapsed
    out = {
        "metric": "field-updates/s",
        "value": value,
        "unit": "field-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"mag_stream: {len(items)} low-field QSC_GRE acquisitions (magLF_spherNeuman), "
                               f"114688 walkers each, 2 fields, one batched launch",
                   "acquisitions": len(items), "field_updates": upd, "walker_steps": walker_steps},
        "batch_ms": float(np.mean(kernel_ms)),
        "create_s": t_create,
        "walker_steps_per_s": walker_steps * args.steps / elapsed,
        "signal_last": [float(v) for v in last.transverse],
    }
    b.close()
    if not args.no_unbatched:
        tu = time.perf_counter()
        for it in items:
            engine.run_mag(*it)
        tun = time.perf_counter() - tu
        out["unbatched_s"] = tun
        out["batched_vs_unbatched"] = tun / (elapsed / args.steps)
    print(json.dumps(out))
    return 0


def bench_batch(args) -> int:
    """The batched uneven workload: 256 instances in one graph launch
    (persistent walk kernel + batched placement and reduction)."""
    import numpy as np

    from axonexa_amd import abi, engine
    from axonexa_amd import workloads as W

    kw = {"scale": args.scale}
    if args.n_steps:
        kw["n_steps"] = args.n_steps
    ws = W.c4_batch(**kw)
    n = ws[0].acq.n_steps
    insts = [(w.sub, w.acq, w.run) for w in ws]
    t0 = time.perf_counter()
    b = engine.Batch(insts, block=args.block, lds_budget=args.lds_budget)
    t_create = time.perf_counter() - t0
    tune = b.autotune(probe_steps=200) if not args.block else {}
    cfg = b.config()
    for _ in range(args.warmup):
        b.run().sync()
    graph_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph_ms.append(b.time())
    elapsed = time.perf_counter() - t0
    total_ws = sum(w.walker_steps for w in ws)
    value = total_ws * args.steps / elapsed
    sigs = [b.results(i) for i in range(len(ws))]
    flops = 0.0
    coll = 0
    for w, sg in zip(ws, sigs):
        flops += w.flops_per_walker_step(sg.collisions / w.walker_steps) * w.walker_steps
        coll += sg.collisions
    avg_ms = float(np.mean(graph_ms))
    achieved = flops / (avg_ms * 1e-3) / 1e12
    out = {
        "metric": "walker-steps/s",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "c4_batch: 128 single spheres R~U(1,8), 64 hex lattices R~U(1,5), "
                        "64 gamma packings; walkers log-uniform 1e4..1e6; 4-32 PGSE/OGSE "
                        "waveforms each; one hipGraph launch",
            "instances": len(ws),
            "walkers": int(sum(w.run.n_walkers for w in ws)),
            "time_steps": n,
            "block": cfg["block"], "walkers_per_thread": cfg["wpt"], "grid": cfg["grid"],
            "tiles": cfg["tiles"], "autotune": tune, "create_s": t_create,
            "lds_budget": args.lds_budget or "default",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None,
            "graph_ms": avg_ms,
            "note": "graph time (placement + persistent walk + reduction) as the duration",
            "collisions_per_walker_step": coll / total_ws,
        },
    }
    if not args.no_unbatched:
        # the same instances one plan at a time (one walk + reduction launch each)
        tu = 0.0
        for w in ws:
            with engine.Plan(w.sub, w.acq, w.run) as p:
                p.sync()
                tp = time.perf_counter()
                p.walk().reduce().sync()
                tu += time.perf_counter() - tp
        out["unbatched"] = {"value": total_ws / tu, "seconds": tu,
                            "speedup_of_batch": value / (total_ws / tu)}
    if not args.no_cpu:
        from oracle import pyoracle as po

        k_per = max(1, args.cpu_walkers // 64)
        threads = _cpu_threads()
        tc = 0.0
        ws_cpu = 0
        bit_all = True
        for i, w in enumerate(ws):
            k = min(k_per, w.run.n_walkers)
            run_k = dataclasses.replace(w.run, n_walkers=k)
            t1 = time.perf_counter()
            ost = po.walk(w.sub, w.acq, run_k, threads=threads)
            tc += time.perf_counter() - t1
            ws_cpu += k * (n - 1)
            gst = b.get_state(i).reshape(-1, w.run.n_walkers)[:, :k].reshape(-1)
            gst = _mask_axes(gst, k, w.acq.n_profiles, po.axes_mask(w.acq))
            bit_all &= bool(np.array_equal(gst, ost))
        out["cpu_baseline"] = {
            "value": ws_cpu / tc, "unit": "walker-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {k_per} walkers of each of the {len(ws)} instances x {n} steps, "
                      f"oracle, {threads} threads on {_cpu_model()}",
            "seconds": tc,
        }
        out["vs_cpu"] = value / (ws_cpu / tc)
        out["signal_error_vs_oracle"] = {"sample_bitwise_equal": bit_all}
    print(json.dumps(out))
    b.close()
    return 0


def bench_surrogate(args) -> int:
    """Learned surrogate: data from the batched engine (Sobol points, gamma
    packings), training with the project's MFMA kernels, held-out error
    against high-walker Monte Carlo, and queries/s of the network at a
    batch of 65,536 against the engine at matched (1 %) noise."""
    import numpy as np

    from axonexa_amd import engine
    from axonexa_amd import surrogate as S

    proto = S.Protocol()
    variants = 8
    G = args.sur_geoms
    Q = G * variants
    u_geo = S.sobol_unit(G, seed=11, dims=3)
    u = np.concatenate([np.repeat(u_geo, variants, axis=0), S.sobol_unit(Q, seed=12, dims=S.N_SCALARS - 3)],
                       axis=1)
    gi = np.repeat(np.arange(G), variants)
    v = S.physical(u)
    t0 = time.perf_counter()
    geoms = S.geometries(u_geo, seed=11, processes=_cpu_threads())
    t_pack = time.perf_counter() - t0
    train_walkers = 2000
    t0 = time.perf_counter()
    Y = S.simulate(v, geoms, gi, proto, train_walkers, seed=1234)
    t_sim = time.perf_counter() - t0
    waves = np.stack([S.waveform_samples(v[q], proto) for q in range(Q)])
    X = S.features(u, waves)
    sur = S.Surrogate(proto)
    sur.init(0)
    t0 = time.perf_counter()
    trace = sur.fit(X, Y.astype(np.float32), steps=args.sur_steps, batch=4096, lr=1e-3)
    t_train = time.perf_counter() - t0
    # held-out geometries (another Sobol scramble), high-walker Monte Carlo as truth
    H = 256
    uh = np.concatenate([S.sobol_unit(H, seed=21, dims=3), S.sobol_unit(H, seed=22, dims=S.N_SCALARS - 3)],
                        axis=1)
    vh = S.physical(uh)
    gh = S.geometries(uh[:, :3], seed=21, processes=_cpu_threads())
    truth_walkers = 100000
    Yh = S.simulate(vh, gh, np.arange(H), proto, truth_walkers, seed=777, first_instance=10 ** 6)
    Xh = S.features(uh, np.stack([S.waveform_samples(vh[q], proto) for q in range(H)]))
    Ph = sur.predict_features(Xh)
    err = np.abs(Ph - Yh)
    # surrogate speed: 65,536 queries resident on the device
    nq = 65536
    Xbig = np.tile(Xh, (nq // H, 1))
    sur.net.forward(Xbig)
    times = [sur.net.time_forward(nq) for _ in range(5)]
    fwd_ms = float(np.median(times))
    sur_qps = nq / (fwd_ms * 1e-3)
    flops = 2.0 * nq * sum(a * b for a, b in zip(sur.dims[:-1], sur.dims[1:]))
    achieved = flops / (fwd_ms * 1e-3) / 1e12
    # Monte Carlo at matched noise: 5000 walkers give sigma ~ 1/sqrt(2N) = 0.01
    mc_walkers = 5000
    nmc = 64
    insts = [S.instance(vh[q], gh[q], proto, mc_walkers, 5, q) for q in range(nmc)]
    with engine.Batch(insts) as b:
        b.run().sync()
        mc_ms = float(np.median([b.time() for _ in range(3)]))
    mc_qps = nmc / (mc_ms * 1e-3)
    # fallback: a query outside the training box goes to Monte Carlo
    uo = uh[:2].copy()
    uo[0, 0] = 1.15          # r_mean beyond the training range
    _, used_mc = sur.predict(uo)
    out = {
        "metric": "surrogate queries/s",
        "value": sur_qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": 1,
        "ms_per_step": fwd_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (signals simulated by this engine)",
        "config": {
            "workload": "c5_surrogate: MLP 268 -> 1024 x 6 (GELU) -> 30, BF16 MFMA, FP32 accumulate",
            "global_batch": nq,
            "train_queries": Q, "train_signals": Q * S.N_DIRS, "train_walkers": train_walkers,
            "adam_steps": args.sur_steps, "batch": 4096,
            "heldout_queries": H, "heldout_walkers": truth_walkers,
            "time_steps": proto.n_steps, "dt_ms": proto.dt,
        },
        "roofline": {
            "bound": "mfma", "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
            "frac": achieved / 2500.0, "traffic": None, "forward_ms": fwd_ms,
        },
        "heldout": {
            "mae": float(err.mean()), "rmse": float(np.sqrt((err ** 2).mean())),
            "p99": float(np.quantile(err, 0.99)), "max": float(err.max()),
            "truth_sigma": float(1.0 / math.sqrt(2 * truth_walkers)),
        },
        "mc": {"queries_per_s": mc_qps, "walkers": mc_walkers, "batch_ms": mc_ms},
        "speedup_vs_mc": sur_qps / mc_qps,
        "fallback_used_mc": [bool(x) for x in used_mc],
        "train": {"loss_first": float(trace[:20].mean()), "loss_last": float(trace[-50:].mean()),
                  "seconds": t_train, "simulate_seconds": t_sim, "pack_seconds": t_pack},
    }
    print(json.dumps(out))
    sur.close()
    return 0


def _mask_axes(st, n_walkers, n_profiles, axes):
    """Zero the moments of axes outside `axes` (the batch keeps every axis,
    the oracle only those in use)."""
    from axonexa_amd import abi

    st = st.copy()
    m = abi.state_view(st, n_walkers, n_profiles)["mom"]
    for a in range(3):
        if not (axes >> a) & 1:
            m[:, :, a] = 0.0
    return st


if __name__ == "__main__":
    sys.exit(main())
