"""Profiling helpers.

    python -m axonexa_amd.profiling copy --mib 1024
        one device-to-device copy of a known size (calibrates the HBM byte
        counters FETCH_SIZE / WRITE_SIZE against a plain copy kernel)

    python -m axonexa_amd.profiling mlp [--batch 65536] [--train-steps 20] [--library]
        the surrogate network (268 -> 1024 x 6 -> 30): 5 forward passes at
        the batch, then Adam steps at batch 4096 on random data (a short
        program for a kernel trace); --library times the same forward pass
        and one 1024 x 1024 layer through torch's bf16 GEMMs instead

    python -m axonexa_amd.profiling summarize <results.db> [--kernel NAME]
        print per-kernel dispatch count, mean duration and the mean of every
        collected counter per dispatch from a rocprofv3 SQLite database
"""
from __future__ import annotations

import argparse
import sqlite3
import sys
from collections import defaultdict


def copy(mib: int) -> None:
    import torch

    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {mib} MiB ({n * 4} bytes read, {n * 4} bytes written)")


def mlp(batch: int, train_steps: int) -> None:
    import numpy as np

    from . import engine

    dims = [268] + [1024] * 6 + [30]
    rng = np.random.default_rng(0)
    with engine.MLP(dims, max_batch=max(batch, 4096)) as m:
        p = np.concatenate([np.concatenate([rng.uniform(-1, 1, a * b) / np.sqrt(a), np.zeros(b)])
                            for a, b in zip(dims[:-1], dims[1:])]).astype(np.float32)
        m.set_params(p)
        x = rng.uniform(-1, 1, size=(batch, 268)).astype(np.float32)
        m.forward(x)
        ms = [m.time_forward(batch) for _ in range(5)]
        flops = 2.0 * batch * sum(a * b for a, b in zip(dims[:-1], dims[1:]))
        print(f"forward batch {batch}: {np.median(ms):.3f} ms, {flops / np.median(ms) / 1e9:.1f} TFLOP/s")
        t = rng.uniform(0, 1, size=(8192, 30)).astype(np.float32)
        m.set_data(x[:8192], t)
        tr = m.train(train_steps, batch=4096, lr=1e-3, seed=1)
        print(f"{train_steps} Adam steps at batch 4096: loss {tr[0]:.4f} -> {tr[-1]:.4f}")


def mlp_library(batch: int) -> None:
    """The same forward pass through torch (its bf16 library GEMMs, bias and
    GELU as separate kernels), timed with events on torch's stream: what the
    vendor library reaches on these shapes.  Run in a process of its own:
    torch's HIP runtime and this package's are separate libraries."""
    import torch

    dims = [268] + [1024] * 6 + [30]
    flops = 2.0 * batch * sum(a * b for a, b in zip(dims[:-1], dims[1:]))
    g = torch.Generator(device="cuda").manual_seed(0)
    ws = [torch.randn(b, a, device="cuda", dtype=torch.bfloat16, generator=g) * a ** -0.5
          for a, b in zip(dims[:-1], dims[1:])]
    bs = [torch.zeros(b, device="cuda", dtype=torch.bfloat16) for b in dims[1:]]
    x = torch.randn(batch, dims[0], device="cuda", dtype=torch.bfloat16, generator=g)

    def run():
        h = x
        for i, (w, b) in enumerate(zip(ws, bs)):
            h = torch.nn.functional.linear(h, w, b)
            if i + 1 < len(ws):
                h = torch.nn.functional.gelu(h, approximate="tanh")
        return h

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        out = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            out.append(e0.elapsed_time(e1))
        return sorted(out)[2]

    ms = timed(run)
    print(f"torch forward (linear + gelu) batch {batch}: {ms:.3f} ms, {flops / ms / 1e9:.1f} TFLOP/s")
    a = torch.randn(batch, 1024, device="cuda", dtype=torch.bfloat16, generator=g)
    w = ws[1].contiguous()
    ms = timed(lambda: torch.matmul(a, w.t()))
    print(f"torch bf16 GEMM {batch} x 1024 x 1024: {ms:.3f} ms, {2.0 * batch * 1024 * 1024 / ms / 1e9:.1f} TFLOP/s")


def _kernels(con):
    rows = con.execute(
        "select k.id, s.string, k.start, k.end from rocpd_kernel_dispatch k "
        "join rocpd_info_kernel_symbol ks on k.kernel_id = ks.id "
        "join rocpd_string s on ks.display_name = s.id").fetchall()
    return rows


def summarize(db: str, kernel: str = "", durations: bool = False) -> str:
    con = sqlite3.connect(db)
    out = []
    if durations:
        # every dispatch of the matching kernels in launch order (short
        # placement dispatches and full walks are told apart here)
        try:
            rows = con.execute("select name, duration from kernels order by start").fetchall()
            out.append(f"{'kernel':<60} {'ms':>12}")
            for name, dur in rows:
                if kernel and kernel not in name:
                    continue
                out.append(f"{name[:60]:<60} {dur / 1e6:>12.4f}")
            out.append("")
        except sqlite3.Error as e:
            out.append(f"(no kernel table: {e})")
    try:
        rows = con.execute("select name, count(*), avg(duration) from kernels group by name").fetchall()
        out.append(f"{'kernel':<60} {'dispatches':>10} {'avg_ms':>12}")
        for name, cnt, dur in sorted(rows, key=lambda r: -r[1] * r[2]):
            if kernel and kernel not in name:
                continue
            out.append(f"{name[:60]:<60} {cnt:>10} {dur / 1e6:>12.4f}")
    except sqlite3.Error as e:
        out.append(f"(no kernel table: {e})")
    try:
        pm = con.execute("select * from counters_collection limit 1")
        cols = [d[0] for d in pm.description]
        q = con.execute("select * from counters_collection").fetchall()
        ci = {c: i for i, c in enumerate(cols)}
        agg = defaultdict(list)
        for r in q:
            name = r[ci.get("kernel_name", ci.get("name", 0))]
            if kernel and kernel not in str(name):
                continue
            cname = r[ci["counter_name"]]
            val = r[ci["value"]]
            disp = r[ci.get("dispatch_id", 0)]
            agg[(str(name), cname)].append((disp, val))
        if agg:
            out.append("")
            out.append(f"{'kernel':<40} {'counter':<28} {'per_dispatch_mean':>20}")
            for (name, cname), vals in sorted(agg.items()):
                per = defaultdict(float)
                for d, v in vals:
                    per[d] += v
                mean = sum(per.values()) / len(per)
                out.append(f"{name[:40]:<40} {cname:<28} {mean:>20.6g}")
    except sqlite3.Error:
        pass
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    ml = sub.add_parser("mlp")
    ml.add_argument("--batch", type=int, default=65536)
    ml.add_argument("--train-steps", type=int, default=20)
    ml.add_argument("--library", action="store_true",
                    help="time the same forward pass through torch's library GEMMs instead")
    s = sub.add_parser("summarize")
    s.add_argument("db")
    s.add_argument("--kernel", default="")
    s.add_argument("--durations", action="store_true", help="also list every dispatch's duration")
    args = ap.parse_args(argv)
    if args.cmd == "copy":
        copy(args.mib)
    elif args.cmd == "mlp":
        mlp_library(args.batch) if args.library else mlp(args.batch, args.train_steps)
    else:
        print(summarize(args.db, args.kernel, args.durations))


if __name__ == "__main__":
    main(sys.argv[1:])
