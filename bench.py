#!/usr/bin/env python3
"""Benchmark of the Klemp-Wilhelmson large step on one MI355X.

Default workload: 400x400x100 cells, dx = dy = 250 m, dz = 200 m, dt = 1.5 s,
8 acoustic substeps, three +2 K bubbles 15 km apart, FP32-storage mixed
precision with hipGraph-captured steps; the plain all-FP64 path is timed on the
same grid for the FP32-mixed / FP64 ratio. A "step" is one large step of the
whole domain with the state resident in HBM. Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2] [--mode explicit|hevi]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s per large step (FP32-mixed vs FP64) + p50/p99 ensemble latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per cell of each kernel, in units of the storage size s
# (4 B mixed, 8 B FP64): every field a kernel must read or write, once.
# The phase totals follow the large-step budget B(N) = (68 + 17 N) s:
# acoustic substep 17 s = forward 13 s (u, v, w, rho, theta, 3 fluxes in;
# rho, p, 3 fluxes out) + backward 4 s of the remaining momentum writes
# (u, v, w written, p read; the inputs shared with the forward pass are not
# counted twice); scalar transport 13 s (9 in + 4 out) split 5 s limiter /
# 8 s update; Kessler 11 s; slow stage 1 + predictor 16 s; stage 2 +
# average + apply 28 s.
ALGO_BYTES_PER_CELL = {
    "acoustic_forward": 13, "acoustic_hevi_column": 13, "acoustic_backward": 4,
    "acoustic_substep": 17,
    "transport_limiter": 5, "transport_update": 8, "transport": 13,
    "kessler_apply": 11,
    "slow_stage1": 16, "slow_stage2": 28,
}


def cpu_info():
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except Exception:
        pass
    return info


def run_member(ctx, L, wl, precision, mode, steps, warmup, graphs):
    dtype = np.float32 if precision == L.PRECISION_MIXED else np.float64
    cfg = wl.config(acoustic_mode=mode)
    m = ctx.member(wl.grid, cfg, wl.base(), precision)
    m.upload(wl.initial_state(dtype))
    ctx.set_graphs(graphs)
    ctx.step([m], warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step([m], steps)
    ctx.synchronize()
    host_s = time.perf_counter() - t0
    dev_ms = ctx.last_step_ms()
    ctx.set_graphs(False)
    return m, dev_ms, host_s


def kernel_table(ctx, m, steps):
    ctx.set_kernel_timing(True)
    ctx.reset_kernel_timing()
    ctx.step([m], steps)
    tab = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    return tab


def cpu_baseline(wl, mode, sample_steps=2):
    from oracle import oracle_py as O
    from nimbus_amd import lib as L

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    g = wl.grid
    og = O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    cfg = O.from_step_config(wl.config(acoustic_mode=mode))
    base = wl.base()
    st = wl.initial_state(np.float32)
    st, _ = O.large_steps(og, base, cfg, st, 1, np.float32)  # warm-up step
    t0 = time.perf_counter()
    _, n_sub = O.large_steps(og, base, cfg, st, sample_steps, np.float32)
    el = time.perf_counter() - t0
    return {"value": wl.cells * sample_steps / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{sample_steps} large steps of the full {wl.describe()} grid after 1 warm-up step, "
                      f"reference storage mode (FP32 storage, FP64 arithmetic), {n_sub} substeps",
            "seconds": el}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--mode", default="explicit", choices=["explicit", "hevi"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="only the mixed member: warm-up + timed steps, for rocprofv3 kernel/counter passes")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run one plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs the large step of one domain per device and "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    from nimbus_amd import lib as L
    from nimbus_amd import workloads as W

    wl = W.NAMED[args.workload]
    mode = L.ACOUSTIC_HEVI if args.mode == "hevi" else L.ACOUSTIC_EXPLICIT
    ctx = L.Context(0)
    graphs = not args.no_graphs

    mixed, mixed_ms, mixed_host = run_member(ctx, L, wl, L.PRECISION_MIXED, mode, args.steps, args.warmup, graphs)
    n_sub = mixed.substeps
    if args.profile:
        print(json.dumps({"profile": True, "workload": wl.describe(), "steps": args.steps,
                          "ms_per_step": mixed_ms / args.steps}))
        mixed.destroy()
        ctx.close()
        return
    tab = kernel_table(ctx, mixed, 2)
    diag = mixed.diag()
    mixed.destroy()
    fp64_steps = max(1, min(args.steps, 10))
    fp64, fp64_ms, _ = run_member(ctx, L, wl, L.PRECISION_FP64, mode, fp64_steps, 1, graphs)
    tab64 = kernel_table(ctx, fp64, 1)
    fp64.destroy()

    ms_per_step = mixed_ms / args.steps
    value = wl.cells * args.steps / (mixed_ms / 1e3)
    fp64_value = wl.cells * fp64_steps / (fp64_ms / 1e3)

    # dominant kernel of the mixed step, with its algorithmic bytes per launch
    total_ms = sum(v[1] for v in tab.values())
    name, (launches, ms) = max(tab.items(), key=lambda kv: kv[1][1])
    avg_ms = ms / launches
    per_cell = ALGO_BYTES_PER_CELL.get(name)
    roofline = None
    if per_cell is not None:
        algo = per_cell * 4 * wl.cells
        achieved = algo / (avg_ms / 1e3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": name,
                    "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": algo,
                    "share_of_step": round(ms / total_ms, 3)}
    step_bytes = (68 + 17 * n_sub) * 4 * wl.cells
    result = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 storage, f64 acoustic/EOS/base-state",
        "data": "synthetic (seeded analytic sounding, warm bubbles, uniform theta noise)",
        "config": {"workload": f"{args.workload}: {wl.describe()}", "cells": wl.cells,
                   "acoustic_mode": args.mode, "substeps": n_sub, "graphs": graphs,
                   "parallelism": "single GPU, one domain"},
        "fp64_value": fp64_value, "fp64_ms_per_step": round(fp64_ms / fp64_steps, 4),
        "mixed_over_fp64": round(value / fp64_value, 3),
        "step_algorithmic_GBps": round(step_bytes / (ms_per_step / 1e3) / 1e9, 1),
        "host_timed_ms_per_step": round(mixed_host * 1e3 / args.steps, 4),
        "kernels_mixed_ms_per_step": {k: round(v[1] / 2, 4) for k, v in sorted(tab.items(), key=lambda kv: -kv[1][1])},
        "kernels_fp64_ms_per_step": {k: round(v[1], 4) for k, v in sorted(tab64.items(), key=lambda kv: -kv[1][1])},
        "diag_after_steps": {k: diag[k] for k in ("max_w", "rms_theta_p", "rain_total_mm")},
        "roofline": roofline,
        "cpu_baseline": None,
        "host": {"platform": platform.platform(), **cpu_info()},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(wl, mode)
        result["gpu_over_cpu"] = round(value / result["cpu_baseline"]["value"], 1)
    ctx.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
