#!/usr/bin/env python3
"""Benchmark of the Klemp-Wilhelmson large step on one MI355X.

Headline workload: 400x400x100 cells, dx = dy = 250 m, dz = 200 m, dt = 1.5 s,
8 acoustic substeps, three +2 K bubbles 15 km apart. `value` is the
cell-updates/s of the mixed-precision path (FP32 bulk in perturbation form,
FP64 base state and FP64 vertically implicit w-p' Thomas solve), hipGraph
captured, state resident in HBM; `fp64_value` is the plain all-FP64 path
(explicit acoustic substeps, the reference method) on the same grid. Also
reported: the other precision/method combinations, the operators beside the
step (the step with explicit diffusion on, member snapshot/restore and NPY
export, CM1 sounding I/O), the 32-member uneven ensemble run as one batch, and
a short continuous-batching serving run with p50/p99 completion latency.
Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]
                  [--serving-seconds S] [--no-ensemble] [--accuracy c2,c1|''] [--no-cpu-baseline]
                  [--no-next-rows]
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

# Algorithmic bytes per cell update of each phase of the large step, in units
# of the storage size s (4 B mixed, 8 B FP64): every field the phase must read
# or write, once (halos, 1-D profiles and 2-D rain free). The budget of one
# large step is B(N) = (68 + 17 N) s for N acoustic substeps:
#   slow stage 1 + predictor 16, stage 2 + average 16, slow apply 12;
#   each acoustic substep 17 (u, v, w, p', rho', theta in; u, v, w, p', rho'
#   out; three retained mass fluxes read and written);
#   scalar transport of 4 scalars 13 (9 in, 4 out);
#   Kessler with sedimentation, apply and EOS 11 (7 in, 4 out).
# A phase is run by the kernels listed with it; `per_step` is how many times
# the phase runs per large step ("substeps" = once per acoustic substep).
# The Smagorinsky SGS and the sponges are outside the budget (reported per
# kernel only).
PHASES = {
    "acoustic_substep": {"bytes_s": 17, "per_step": "substeps",
                         "kernels": ("acoustic_forward", "acoustic_hevi_explicit", "acoustic_hevi_column",
                                     "acoustic_hevi_column_seg", "acoustic_momentum", "acoustic_commit",
                                     "bc_west_east", "bc_south_north", "bc_vertical", "seed_radiation")},
    "slow_rk2_apply": {"bytes_s": 44, "per_step": 1, "kernels": ("slow_stage1", "slow_stage2")},
    "scalar_transport": {"bytes_s": 13, "per_step": 1,
                         "kernels": ("transport_flux", "transport_limiter", "transport_update")},
    "kessler": {"bytes_s": 11, "per_step": 1, "kernels": ("kessler_apply", "scalars_home")},
}


def phase_table(tab, cells, elem, n_sub, steps, traffic=None):
    """Per phase: the budget bytes of one execution, the summed duration of its
    kernels per execution (HIP events on the launch stream, `steps` large
    steps), achieved GB/s and fraction of the HBM peak; `traffic` (kernel ->
    counted HBM bytes per launch) adds the measured bytes per execution."""
    out = {}
    for name, ph in PHASES.items():
        per_step = n_sub if ph["per_step"] == "substeps" else ph["per_step"]
        execs = per_step * steps
        ms = sum(tab[k][1] for k in ph["kernels"] if k in tab)
        if ms <= 0:
            continue
        algo = ph["bytes_s"] * elem * cells
        t = ms / execs
        gbs = algo / (t / 1e3) / 1e9
        row = {"kernels": {k: {"launches_per_execution": tab[k][0] / execs,
                               "avg_launch_ms": round(tab[k][1] / tab[k][0], 4)}
                           for k in ph["kernels"] if k in tab},
               "bytes_per_cell": ph["bytes_s"] * elem, "algorithmic_bytes_per_execution": algo,
               "ms_per_execution": round(t, 4), "executions_per_step": per_step,
               "achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        if traffic is not None:
            have = [k for k in row["kernels"] if k in traffic]
            if have:
                counted = sum(traffic[k] * row["kernels"][k]["launches_per_execution"] for k in have)
                row["traffic_bytes_per_execution"] = int(counted)
                row["traffic_over_budget"] = round(counted / algo, 2)
                row["traffic_kernels_without_counters"] = [k for k in row["kernels"] if k not in traffic]
        out[name] = row
    return out


# Bytes each kernel itself reads and writes once per cell (its own I/O, which
# exceeds the phase budget where a phase is split over several kernels that
# pass intermediates through memory). In units of s, plus FP64 bytes.
ALGO_PER_CELL = {
    # read u v w rho' theta mfx mfy mfz; write rho'(new) p' mfx mfy mfz
    "acoustic_forward": (13, 0),
    # read rho'(new) p' u v w; write u v, and w for explicit members
    "acoustic_momentum": {"explicit": (8, 0), "hevi": (7, 0)},
    # HE-VI horizontal part: read rho' u v theta mfx mfy; write mfx mfy rho* p*
    "acoustic_hevi_explicit": (10, 0),
    # HE-VI column solve: read rho* p* p' w rho' theta mfz; write w rho' p' mfz;
    # the FP64 Thomas coefficients c', d' are written and read back (32 B)
    "acoustic_hevi_column": (11, 32),
    # the partitioned column solve of FP64 members: the same 11 fields, no coefficient arrays
    "acoustic_hevi_column_seg": (11, 0),
    # read mfx mfy mfz and four scalars; write 3 face fluxes per scalar
    "transport_flux": (19, 0),
    # read rho_b, and fluxes and value of the three moisture scalars; write 3 limiter ratios
    "transport_limiter": (16, 0),
    # read rho_b rho_e, 12 fluxes, 3 ratios, 4 scalars; write 4 scalars
    "transport_update": (25, 0),
    # read theta qv qc qr rho' p'; write theta qv qc qr p'
    "kessler_apply": (11, 0),
    # read u v w rho' qv qc qr; write k1 (3) and the predictor (3)
    "slow_stage1": (13, 0),
    # read predictor (3) rho' qv qc qr, k1 (3), u v w; write u v w
    "slow_stage2": (16, 0),
    # Smagorinsky substep: read u v w; write nu K_theta K_q
    "sgs_closure": (6, 0),
    # read u v w rho' nu; write u v w (the other buffer set)
    "sgs_momentum": (8, 0),
    # read theta qv rho' K_theta K_q; write theta qv (p' of the last substep not counted)
    "sgs_scalars": (7, 0),
    # explicit Laplacian diffusion of u v w theta qv: read the five fields and
    # rho'; write the five tendencies
    "diffusion_tendencies": (11, 0),
    # read u v w theta qv, the five tendencies and rho'; write u v w theta qv p'
    "diffusion_commit": (17, 0),
}


_T0 = time.perf_counter()


def progress(msg):
    """One line per phase on stderr (stdout carries only the result line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def algo_bytes_per_cell(name, mode, elem):
    e = ALGO_PER_CELL.get(name)
    if isinstance(e, dict):
        e = e.get(mode)
    if e is None:
        return None
    return e[0] * elem + e[1]


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


# explicit Laplacian diffusion timed beside the step (its coefficients are zero
# in the reference's presets): u, v, w, theta and qv, K_h 50 and K_v 10 m^2/s
DIFFUSION_TIMED = {"apply": "all", "kh": 50.0, "kv": 10.0}


def run_member(ctx, L, wl, precision, mode, steps, warmup, graphs, column_solve=None, diffusion=None):
    dtype = np.float32 if precision == L.PRECISION_MIXED else np.float64
    cfg = wl.config(acoustic_mode=mode, diffusion=diffusion)
    m = ctx.member(wl.grid, cfg, wl.base(), precision)
    if column_solve is not None:
        m.set_column_solve(column_solve)
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


def _oracle_threads():
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    return threads


def cpu_baseline(wl, sample_steps=10, warmup=2):
    """The oracle in the reference's CPU arithmetic (FP32 storage, FP64
    arithmetic, explicit acoustic substeps) on the host cores: `warmup` large
    steps, then `sample_steps` timed; and one step on one thread."""
    from oracle import oracle_py as O

    threads = _oracle_threads()
    g = wl.grid
    og = O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    cfg = O.from_step_config(wl.config())
    base = wl.base()
    st = wl.initial_state(np.float32)
    st, _ = O.large_steps(og, base, cfg, st, warmup, np.float32)
    t0 = time.perf_counter()
    st2, n_sub = O.large_steps(og, base, cfg, st, sample_steps, np.float32)
    el = time.perf_counter() - t0
    # the same arithmetic on one thread: one large step of the full grid
    prev = O.set_threads(1)
    t1 = time.perf_counter()
    O.large_steps(og, base, cfg, st2, 1, np.float32)
    el1 = time.perf_counter() - t1
    O.set_threads(prev)
    return {"value": wl.cells * sample_steps / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{sample_steps} large steps of the full {wl.describe()} grid after {warmup} warm-up steps, "
                      f"reference arithmetic (FP32 storage, FP64 arithmetic, explicit substeps), {n_sub} substeps",
            "seconds": round(el, 3),
            "one_hour_extrapolated_s": round(el / sample_steps * 3600.0 / wl.dt, 1),
            "single_thread": {"value": wl.cells / el1, "unit": "cell-updates/s", "cores": 1,
                              "sample": "1 large step of the same grid and state", "seconds": round(el1, 3)}}


def cpu_baseline_more(wl, W, b_steps=4):
    """More of the oracle on the host cores: the FP64 B mode (FP64 storage,
    HE-VI) on the same grid, the 64 x 64 x 40 plain FP64 method over its 10
    steps, and the 32-member ensemble (each member 1 warm-up + 2 timed steps,
    the rates combined as total cell updates over total time)."""
    from oracle import oracle_py as O

    threads = _oracle_threads()
    out = {}
    g = wl.grid
    og = O.grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz)
    import nimbus_amd.lib as L
    cfg = O.from_step_config(wl.config(acoustic_mode=L.ACOUSTIC_HEVI))
    st = wl.initial_state(np.float64)
    st, _ = O.large_steps(og, wl.base(), cfg, st, 1, np.float64)
    t0 = time.perf_counter()
    _, n_sub = O.large_steps(og, wl.base(), cfg, st, b_steps, np.float64)
    el = time.perf_counter() - t0
    out["b_mode_fp64_hevi"] = {"value": wl.cells * b_steps / el, "unit": "cell-updates/s", "cores": threads,
                               "kind": "port", "seconds": round(el, 3),
                               "sample": f"{b_steps} FP64 HE-VI large steps of {wl.describe()} after 1 warm-up "
                                         f"step, {n_sub} substeps"}
    c0 = W.NAMED["c0"]
    g0 = c0.grid
    cfg0 = O.from_step_config(c0.config())
    t0 = time.perf_counter()
    _, n0 = O.large_steps(O.grid(g0.nx, g0.ny, g0.nz, g0.dx, g0.dy, g0.dz), c0.base(), cfg0,
                          c0.initial_state(np.float64), c0.steps, np.float64)
    el0 = time.perf_counter() - t0
    out["c0_plain_fp64"] = {"value": c0.cells * c0.steps / el0, "unit": "cell-updates/s", "cores": threads,
                            "kind": "port", "seconds": round(el0, 3),
                            "sample": f"all {c0.steps} steps of {c0.describe()}, FP64, explicit, {n0} substeps"}
    cells_steps, secs = 0, 0.0
    for m in W.ensemble_c3():
        gm = m.grid
        ogm = O.grid(gm.nx, gm.ny, gm.nz, gm.dx, gm.dy, gm.dz)
        cm = O.from_step_config(m.config())
        s, _ = O.large_steps(ogm, m.base(), cm, m.initial_state(np.float32), 1, np.float32)
        t0 = time.perf_counter()
        O.large_steps(ogm, m.base(), cm, s, 2, np.float32)
        secs += time.perf_counter() - t0
        cells_steps += 2 * m.cells
    out["ensemble_c3"] = {"value": cells_steps / secs, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                          "seconds": round(secs, 3),
                          "sample": "each of the 32 members: 1 warm-up + 2 timed steps in the reference "
                                    "arithmetic (FP32 storage, FP64 arithmetic, explicit substeps), total cell "
                                    "updates over total time"}
    return out


def measured_traffic_all(workload, mode):
    """kernel -> counted HBM bytes per launch from the committed counter summary."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}_{mode}.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    return {k: int(v["fetch_kib"] * 1024 * d["fetch_scale"] + v["write_kib"] * 1024 * d["write_scale"])
            for k, v in d.get("kernels", {}).items()}


def roofline_for(tab, cells, elem, workload, mode, n_sub, steps):
    """The dominant phase of the step (largest summed kernel time): its budget
    bytes per execution / the summed duration of its kernels per execution."""
    phases = phase_table(tab, cells, elem, n_sub, steps, measured_traffic_all(workload, mode))
    name, ph = max(phases.items(), key=lambda kv: kv[1]["ms_per_execution"] * kv[1]["executions_per_step"])
    total_ms = sum(v[1] for v in tab.values()) / steps
    return {"bound": "hbm", "achieved": ph["achieved_GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ph["frac"], "traffic": ph.get("traffic_bytes_per_execution"),
            "traffic_over_budget": ph.get("traffic_over_budget"), "phase": name,
            "kernels": ph["kernels"], "avg_execution_ms": ph["ms_per_execution"],
            "algorithmic_bytes_per_execution": ph["algorithmic_bytes_per_execution"],
            "bytes_per_cell": ph["bytes_per_cell"],
            "share_of_step": round(ph["ms_per_execution"] * ph["executions_per_step"] / total_ms, 3)}


def column_solve_code(L, name):
    return {"auto": None, "sequential": L.COLUMN_SOLVE_SEQUENTIAL, "partitioned": L.COLUMN_SOLVE_PARTITIONED}[name]


def ensemble_run(ctx, L, W, precision, mode, tuning=None, column_solve=None):
    """The 32-member uneven ensemble as one resident batch, 30 simulated minutes
    per member; members leave the batch as they finish."""
    from nimbus_amd import serving as SV
    reqs = [SV.Request(q, 0.0, wl) for q, wl in enumerate(W.ensemble_c3())]
    before = ctx.batch_stats()
    srv = SV.Server(ctx, budget_bytes=250e9, precision=precision, acoustic_mode=mode, quantum=50, tuning=tuning,
                    column_solve=column_solve)
    rep = srv.run(reqs)
    s = rep.summary()
    after = ctx.batch_stats()
    lat = sorted(r.latency for r in rep.results)
    return {"members": len(reqs), "cells_total": sum(r.workload.cells for r in reqs),
            "cell_updates_per_s": s["cell_updates_per_s"], "wall_seconds": s["wall_seconds"],
            "device_step_seconds": s["device_step_seconds"], "p50_completion_s": s["p50_latency_s"],
            "p99_completion_s": s["p99_latency_s"], "max_completion_s": round(lat[-1], 3),
            "rounds": s["rounds"], "acoustic_mode": "hevi" if mode == L.ACOUSTIC_HEVI else "explicit",
            "column_solve": {None: "per-member default", L.COLUMN_SOLVE_SEQUENTIAL: "sequential",
                             L.COLUMN_SOLVE_PARTITIONED: "partitioned"}[column_solve],
            "autotuned_shapes": tuning.tuned if tuning else 0,
            "membership_changes": {k: round(after[k] - before[k], 3) for k in after},
            "data": "seed 3: nx = ny in {48..160}, nz in {40, 60, 80}, dx in {500, 1000} m, "
                    "dt = 6 min(dx, dz)/1000 s, 6 substeps, perturbed soundings, 1800 simulated s each"}


def serving_run(ctx, L, seconds, precision, mode, fp64_sample=5, tuning=None, column_solve=None,
                stop_after=None, max_cells=None, admission="fifo"):
    """Continuous batching of seeded Poisson requests (2/s) arriving over `seconds`,
    then `fp64_sample` requests (drawn with seed 5) rerun alone in FP64: the error of each
    served (mixed) result against its FP64 rerun. `stop_after` (wall seconds) ends
    the run early and reports the backlog left."""
    import random
    from nimbus_amd import serving as SV
    from nimbus_amd import accuracy as ACC
    reqs = SV.poisson_requests(seed=4, rate=2.0, horizon=float(seconds))
    sample = sorted(random.Random(5).sample([r.rid for r in reqs], min(fp64_sample, len(reqs))))
    before = ctx.batch_stats()
    srv = SV.Server(ctx, budget_bytes=200e9, precision=precision, acoustic_mode=mode, quantum=16, tuning=tuning,
                    column_solve=column_solve, progress=lambda m: progress("  serving " + m),
                    max_resident_cells=max_cells, admission=admission)
    rep = srv.run(reqs, keep_fields=sample, stop_after=stop_after)
    s = rep.summary()
    after = ctx.batch_stats()
    s["arrival_seconds"] = seconds
    s["stream"] = "full 10-minute stream" if seconds >= 600 else f"reduced stream: the first {seconds:g} s of arrivals"
    if stop_after is not None:
        s["stream"] += (f", stopped after {stop_after:g} s of wall time with {s['unfinished']} requests unfinished"
                        if s["unfinished"] else f", finished within the {stop_after:g} s stop")
    s["requests_generated"] = len(reqs)
    s["max_resident_cells"] = max_cells
    s["admission"] = admission
    s["membership_changes"] = {k: round(after[k] - before[k], 3) for k in after}
    s["data"] = ("seed 4 Poisson arrivals at 2/s; nx = ny in {48..256}, nz in {40, 60, 80}, dx 500-1000 m, "
                 "15-90 simulated minutes each; 200 GB HBM budget")
    by_id = {r.rid: r for r in reqs}
    errors = []
    for res in rep.results:
        if res.fields is None:
            continue
        req = by_id[res.rid]
        progress(f"  FP64 rerun of request {res.rid}: {req.workload.describe()}, {req.workload.steps} steps")
        fields, diag = SV.run_alone(ctx, req, precision=L.PRECISION_FP64, acoustic_mode=mode)
        e = ACC.compare(req.workload, res.fields, fields, res.diag, diag)
        errors.append({"rid": res.rid, "grid": req.workload.describe(), "steps": req.workload.steps,
                       "max_w_error": e["max_w_error"], "rms_theta_p_error": e["rms_theta_p_error"],
                       "rain_relative_error": e["rain_relative_error"], "targets_met": e["targets_met"]})
    s["fp64_rerun"] = errors
    return s


def next_rows_run(ctx, L, wl, steps, warmup, graphs, base_ms):
    """The operators beside the large step, at the same workload: the mixed
    HE-VI step with the explicit Laplacian diffusion of u, v, w, theta and qv on
    (K_h 50, K_v 10 m^2/s) and its two kernels against their own reads and
    writes; the member snapshot (device to host), restore (host to device) and
    NPY field export; the CM1 sounding text written, read and turned into the
    base state of the grid."""
    import shutil
    import tempfile
    out = {}
    m, ms, _ = run_member(ctx, L, wl, L.PRECISION_MIXED, L.ACOUSTIC_HEVI, steps, warmup, graphs,
                          diffusion=DIFFUSION_TIMED)
    ms /= steps
    tab = kernel_table(ctx, m, 2)
    kern = {}
    for name in ("diffusion_tendencies", "diffusion_commit"):
        if name not in tab or tab[name][0] == 0:
            continue
        launches, kms = tab[name]
        per_cell = algo_bytes_per_cell(name, "hevi", 4)
        gbs = per_cell * wl.cells / (kms / launches / 1e3) / 1e9
        kern[name] = {"launches_per_step": launches / 2, "avg_launch_ms": round(kms / launches, 4),
                      "bytes_per_cell": per_cell, "achieved_GBps": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["diffusion"] = {"workload": wl.describe(),
                        "coefficients": {"K_h": DIFFUSION_TIMED["kh"], "K_v": DIFFUSION_TIMED["kv"]},
                        "fields": "u v w theta qv", "ms_per_step": round(ms, 4),
                        "ms_per_step_without": round(base_ms, 4),
                        "cell_updates_per_s": wl.cells / (ms / 1e3), "kernels": kern}

    t0 = time.perf_counter()
    blob = m.snapshot()
    t_snap = time.perf_counter() - t0
    t0 = time.perf_counter()
    m2 = ctx.restore(blob)
    t_rest = time.perf_counter() - t0
    same = m2.snapshot() == blob
    m2.destroy()
    d = tempfile.mkdtemp(prefix="nimbus_npy_")
    try:
        t0 = time.perf_counter()
        m.export_npy(d)
        t_npy = time.perf_counter() - t0
        npy_bytes = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    m.destroy()
    out["state_io"] = {"snapshot_bytes": len(blob), "snapshot_s": round(t_snap, 4),
                       "snapshot_GBps": round(len(blob) / t_snap / 1e9, 2), "restore_s": round(t_rest, 4),
                       "restore_GBps": round(len(blob) / t_rest / 1e9, 2), "restored_snapshot_identical": same,
                       "npy_export_bytes": npy_bytes, "npy_export_s": round(t_npy, 4),
                       "npy_export_GBps": round(npy_bytes / t_npy / 1e9, 2),
                       "note": "host wall time; snapshot into pageable host memory, NPY files written to /tmp"}

    g = wl.grid
    z = np.concatenate([[0.0], (np.arange(g.nz) + 0.5) * g.dz, [g.nz * g.dz]])
    levels = L.sounding_levels(z, wl.sounding_struct())
    reps = 200
    t0 = time.perf_counter()
    for _ in range(reps):
        text = L.cm1_format(levels)
    t_fmt = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        col = L.cm1_parse(text)
    t_parse = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        L.cm1_base_state(wl.grid, col)
    t_base = (time.perf_counter() - t0) / reps
    out["sounding_io"] = {"levels": len(col["z"]), "text_bytes": len(text), "write_ms": round(t_fmt * 1e3, 4),
                          "read_ms": round(t_parse * 1e3, 4), "base_state_ms": round(t_base * 1e3, 4),
                          "note": f"host, mean of {reps} calls through the C ABI"}
    out["reference_contracts"] = contract_rates(ctx, L, wl)
    return out


def contract_rates(ctx, L, wl, reps=2):
    """The reference's per-operation dispatch contracts take host arrays, so a
    call moves its inputs to the device and its outputs back: wall time per
    call at the workload's grid (pageable numpy arrays, after one warm-up
    call) for the conservative transport of four scalars with supplied mass
    fluxes and the explicit diffusion of five fields, with the bytes that
    cross the host link and the resulting cell-updates/s."""
    import ctypes
    g = wl.grid
    st = wl.initial_state(np.float32)
    shape = (g.nx, g.ny, g.nz)
    rho = np.ascontiguousarray(st["rho"], dtype=np.float32)
    fl = {k: np.ascontiguousarray(st[k], dtype=np.float32) for k in ("u", "v", "w", "theta", "qv", "qc", "qr")}
    mf = [np.ascontiguousarray(rho * fl[k]) for k in ("u", "v", "w")]
    mf[0][-1] = 0.0
    mf[1][:, -1] = 0.0
    mf[2][:, :, -1] = 0.0

    def timed(fn, call, nbytes):
        call()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        sec = (time.perf_counter() - t0) / reps
        return {"entry_point": fn, "ms_per_call": round(sec * 1e3, 3), "host_link_bytes": int(nbytes),
                "host_link_GBps": round(nbytes / sec / 1e9, 2), "cell_updates_per_s": wl.cells / sec}

    out = {}
    qs = [fl[k] for k in ("theta", "qv", "qc", "qr")]
    outs = [np.empty(shape, np.float32) for _ in qs]
    qin = (ctypes.POINTER(ctypes.c_float) * 4)(*[L.fptr(q) for q in qs])
    qout = (ctypes.POINTER(ctypes.c_float) * 4)(*[L.fptr(o) for o in outs])
    pd = (ctypes.c_int * 4)(0, 1, 1, 1)
    kap = np.zeros(4, np.float32)
    d = L.ConservativeTransportDispatch(None, None, None, L.fptr(rho), qin, qout, L.fptr(kap), pd, 4, g.nx, g.ny,
                                        g.nz, 1.0 / g.dx, 1.0 / g.dy, 1.0 / g.dz, 0.0, 0.0, 0.0, wl.dt, 0.05,
                                        L.fptr(rho), L.fptr(mf[0]), L.fptr(mf[1]), L.fptr(mf[2]), 1, 0.0, 0.0, 1, 0)
    fn = "nimbus_hip_dispatch_advect_scalars_conservative_cartesian_cgrid"
    entry = getattr(ctx.lib, fn)
    out["conservative_transport_4_scalars"] = timed(
        fn, lambda: L.check(entry(ctx.handle, ctypes.byref(d)), fn), 4 * wl.cells * (9 + 4))
    names = ("u", "v", "w", "theta", "qv")
    locs = (L.DIFFUSION_U_FACE, L.DIFFUSION_V_FACE, L.DIFFUSION_W_FACE, L.DIFFUSION_SCALAR, L.DIFFUSION_SCALAR)
    tends = [np.empty(shape, np.float32) for _ in names]
    dz = np.full(g.nz, g.dz, np.float32)
    fields = (ctypes.POINTER(ctypes.c_float) * 5)(*[L.fptr(fl[k]) for k in names])
    tarr = (ctypes.POINTER(ctypes.c_float) * 5)(*[L.fptr(t) for t in tends])
    larr = (ctypes.c_int * 5)(*locs)
    dd = L.DiffusionBatchDispatch(fields, tarr, larr, 5, L.fptr(rho), L.fptr(dz), g.nx, g.ny, g.nz, g.dx, g.dy,
                                  DIFFUSION_TIMED["kh"], DIFFUSION_TIMED["kv"], 1, 0, 0, None)
    fn = "nimbus_hip_dispatch_diffusion_batch"
    entry2 = getattr(ctx.lib, fn)
    out["diffusion_batch_5_fields"] = timed(
        fn, lambda: L.check(entry2(ctx.handle, ctypes.byref(dd)), fn), 4 * wl.cells * (6 + 5))
    out["note"] = ("host arrays in and out on every call, as the reference's dispatch contracts pass them; "
                   "the resident step runs the same operators in 1.1 ms (transport) and 0.8 ms (diffusion)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--serving-seconds", type=float, default=20.0,
                    help="arrival window of the serving run (600 = the full 10-minute stream; 0 skips it)")
    ap.add_argument("--serving-stop-after", type=float, default=None,
                    help="end the serving run after this many wall seconds and report the backlog")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving run (with --serving-seconds), no single-member or ensemble runs")
    ap.add_argument("--serving-admission", choices=["fifo", "shortest"], default="fifo",
                    help="order of the waiting requests: arrival, or least work first")
    ap.add_argument("--serving-max-cells", type=float, default=48e6,
                    help="cap on the cells resident at once in the serving run besides the 200 GB HBM budget "
                         "(0: no cap)")
    ap.add_argument("--column-solve", choices=["auto", "sequential", "partitioned"], default="sequential",
                    help="HE-VI column-solve form of the ensemble and serving members (FP32 storage)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="ensemble and serving members keep the default launch geometry (no autotuning)")
    ap.add_argument("--accuracy", default="c2,c1",
                    help="workloads of the 1-hour mixed-vs-FP64 error runs, comma-separated; the first is "
                         "reported as accuracy_1h ('' skips them)")
    ap.add_argument("--profile", choices=["hevi", "explicit"], default=None,
                    help="only one member in this mode: warm-up + timed steps, for rocprofv3 passes")
    ap.add_argument("--profile-precision", choices=["mixed", "fp64"], default="mixed",
                    help="storage of the --profile member")
    ap.add_argument("--profile-diffusion", action="store_true",
                    help="with --profile: the explicit diffusion of u, v, w, theta, qv on (K_h 50, K_v 10 m^2/s)")
    ap.add_argument("--profile-column-solve", choices=["auto", "sequential", "partitioned"], default="auto",
                    help="HE-VI column-solve form of the --profile member (auto: the member's default)")
    ap.add_argument("--no-next-rows", action="store_true",
                    help="skip the diffusion, state I/O and sounding I/O timings")
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
        print(f"bench.py measures one GPU: the large step runs on one device and "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    from nimbus_amd import lib as L
    from nimbus_amd import workloads as W

    wl = W.NAMED[args.workload]
    ctx = L.Context(0)
    graphs = not args.no_graphs
    modes = {"hevi": L.ACOUSTIC_HEVI, "explicit": L.ACOUSTIC_EXPLICIT}
    tuning = None
    if not args.no_tuning:
        from nimbus_amd import tuning as T
        tuning = T.TuningCache(ctx.device_arch())
    col = column_solve_code(L, args.column_solve)

    if args.serving_only:
        progress(f"serving: Poisson arrivals over {args.serving_seconds:g} s")
        out = {"serving_c4": serving_run(ctx, L, args.serving_seconds, L.PRECISION_MIXED, L.ACOUSTIC_HEVI,
                                         tuning=tuning, column_solve=col, stop_after=args.serving_stop_after,
                                         fp64_sample=0, max_cells=args.serving_max_cells or None,
                                         admission=args.serving_admission)}
        ctx.close()
        print(json.dumps(out))
        return

    if args.profile:
        prec = L.PRECISION_FP64 if args.profile_precision == "fp64" else L.PRECISION_MIXED
        m, ms, _ = run_member(ctx, L, wl, prec, modes[args.profile], args.steps, args.warmup, graphs,
                              column_solve_code(L, args.profile_column_solve),
                              DIFFUSION_TIMED if args.profile_diffusion else None)
        print(json.dumps({"profile": True, "workload": wl.describe(), "mode": args.profile,
                          "precision": args.profile_precision, "column_solve": m.column_solve(),
                          "diffusion": DIFFUSION_TIMED if args.profile_diffusion else None,
                          "steps": args.steps, "ms_per_step": ms / args.steps}))
        m.destroy()
        ctx.close()
        return

    runs = {}
    tables = {}
    for label, prec, mode in (("mixed_hevi", L.PRECISION_MIXED, "hevi"),
                              ("mixed_explicit", L.PRECISION_MIXED, "explicit"),
                              ("fp64_explicit", L.PRECISION_FP64, "explicit"),
                              ("fp64_hevi", L.PRECISION_FP64, "hevi")):
        progress(f"{label}: {wl.describe()}")
        steps = args.steps if prec == L.PRECISION_MIXED else max(1, min(args.steps, 10))
        m, ms, _ = run_member(ctx, L, wl, prec, modes[mode], steps, args.warmup if prec == L.PRECISION_MIXED else 1,
                              graphs)
        runs[label] = {"ms_per_step": ms / steps, "value": wl.cells * steps / (ms / 1e3), "substeps": m.substeps}
        tables[label] = kernel_table(ctx, m, 2)
        if label == "mixed_hevi":
            diag = m.diag()
        m.destroy()

    head = runs["mixed_hevi"]
    value = head["value"]
    fp64_value = runs["fp64_explicit"]["value"]
    n_sub = head["substeps"]
    ms_per_step = head["ms_per_step"]
    # algorithmic bytes of one mixed HE-VI step: the table above summed over the
    # kernels that ran (kernel tables cover 2 steps)
    step_bytes = sum(launches / 2 * algo_bytes_per_cell(name, "hevi", 4) * wl.cells
                     for name, (launches, _) in tables["mixed_hevi"].items()
                     if algo_bytes_per_cell(name, "hevi", 4) is not None)

    def per_step(tab):
        return {k: round(v[1] / 2, 4) for k, v in sorted(tab.items(), key=lambda kv: -kv[1][1])}

    def per_kernel_roofline(tab, mode):
        """Every kernel of a step with a counted byte figure: algorithmic GB/s
        of its average launch and the fraction of the HBM peak."""
        out = {}
        for name, (launches, ms) in sorted(tab.items(), key=lambda kv: -kv[1][1]):
            per_cell = algo_bytes_per_cell(name, mode, 4)
            if per_cell is None or launches == 0 or ms <= 0:
                continue
            gbs = per_cell * wl.cells / (ms / launches / 1e3) / 1e9
            out[name] = {"bytes_per_cell": per_cell, "avg_launch_ms": round(ms / launches, 4),
                         "achieved_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        return out

    result = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 storage and bulk arithmetic; f64 base state, continuity/EOS, vertical Thomas solve",
        "data": "synthetic (seeded analytic sounding, warm bubbles, uniform theta noise)",
        "config": {"workload": f"{args.workload}: {wl.describe()}", "cells": wl.cells,
                   "acoustic_mode": "hevi (vertically implicit, FP64 Thomas)", "substeps": n_sub,
                   "graphs": graphs, "parallelism": "single GPU, one domain"},
        "fp64_value": fp64_value, "fp64_ms_per_step": round(runs["fp64_explicit"]["ms_per_step"], 4),
        "mixed_over_fp64": round(value / fp64_value, 3),
        "modes": {k: {"cell_updates_per_s": v["value"], "ms_per_step": round(v["ms_per_step"], 4),
                      "substeps": v["substeps"]} for k, v in runs.items()},
        "mixed_over_fp64_same_method": {
            "hevi": round(runs["mixed_hevi"]["value"] / runs["fp64_hevi"]["value"], 3),
            "explicit": round(runs["mixed_explicit"]["value"] / runs["fp64_explicit"]["value"], 3)},
        "step_algorithmic_GBps": round(step_bytes / (ms_per_step / 1e3) / 1e9, 1),
        "kernels_ms_per_step": {k: per_step(t) for k, t in tables.items()},
        "diag_after_steps": {k: diag[k] for k in ("max_w", "rms_theta_p", "rain_total_mm")},
        "roofline": roofline_for(tables["mixed_hevi"], wl.cells, 4, args.workload, "hevi", n_sub, 2),
        # whole step against the large-step byte budget B(N) = (68 + 17 N) s per cell update
        "roofline_step_budget": {"bytes_per_cell_update": (68 + 17 * n_sub) * 4,
                                 "achieved": round(value * (68 + 17 * n_sub) * 4 / 1e9, 1),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(value * (68 + 17 * n_sub) * 4 / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline_mixed_explicit": roofline_for(tables["mixed_explicit"], wl.cells, 4, args.workload, "explicit",
                                                runs["mixed_explicit"]["substeps"], 2),
        "phases_roofline": {k: phase_table(tables[k], wl.cells, 8 if k.startswith("fp64") else 4,
                                           runs[k]["substeps"], 2) for k in tables},
        "kernels_io_roofline": {"mixed_hevi": per_kernel_roofline(tables["mixed_hevi"], "hevi"),
                                "mixed_explicit": per_kernel_roofline(tables["mixed_explicit"], "explicit")},
        "cpu_baseline": None,
        "host": {"platform": platform.platform(), **cpu_info()},
    }
    if not args.no_next_rows:
        progress("beside the step: diffusion, state I/O, sounding I/O")
        result["next_rows"] = next_rows_run(ctx, L, wl, min(args.steps, 10), args.warmup, graphs, ms_per_step)
    if not args.no_ensemble:
        progress("ensemble: 32 members, 30 simulated minutes each")
        result["ensemble_c3"] = ensemble_run(ctx, L, W, L.PRECISION_MIXED, L.ACOUSTIC_HEVI, tuning, col)
    if args.serving_seconds > 0:
        progress(f"serving: Poisson arrivals over {args.serving_seconds:g} s")
        result["serving_c4"] = serving_run(ctx, L, args.serving_seconds, L.PRECISION_MIXED, L.ACOUSTIC_HEVI,
                                           tuning=tuning, column_solve=col, stop_after=args.serving_stop_after,
                                           max_cells=args.serving_max_cells or None,
                                           admission=args.serving_admission)
    if args.accuracy:
        from nimbus_amd import accuracy as ACC
        for q, name in enumerate(n for n in args.accuracy.split(",") if n):
            progress(f"accuracy: one simulated hour of {name}, mixed / FP64 / FP64 control")
            acc = ACC.one_hour(ctx, W.NAMED[name], L.ACOUSTIC_HEVI,
                               progress=lambda t, name=name: progress(f"  {name}: {t:g} simulated s"))
            result["accuracy_1h" if q == 0 else f"accuracy_1h_{name}"] = acc
    ctx.close()
    if not args.no_cpu_baseline:
        progress("cpu baseline: the oracle on the host cores")
        result["cpu_baseline"] = cpu_baseline(wl)
        result["gpu_over_cpu"] = round(value / result["cpu_baseline"]["value"], 1)
        progress("cpu baseline: FP64 B mode, the 64x64x40 plain method, the ensemble")
        result["cpu_baseline_more"] = cpu_baseline_more(wl, W)
        result["gpu_over_cpu_more"] = {
            "fp64_hevi": round(runs["fp64_hevi"]["value"] / result["cpu_baseline_more"]["b_mode_fp64_hevi"]["value"],
                               1)}
        if "ensemble_c3" in result:
            result["gpu_over_cpu_more"]["ensemble"] = round(
                result["ensemble_c3"]["cell_updates_per_s"] / result["cpu_baseline_more"]["ensemble_c3"]["value"], 1)
    progress("done")
    print(json.dumps(result))


if __name__ == "__main__":
    main()
