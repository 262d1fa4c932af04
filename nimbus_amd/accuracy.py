"""Error of the mixed-precision path against the all-FP64 path after a long run.

The FP64 device path is the reference here: it agrees with the CPU FP64 oracle
to ~1e-9 of each field's scale over several large steps, and a simulated hour
of the 160x160x80 or 400x400x100 workloads is out of the CPU oracle's reach
within a run. Both runs start from the same seeded state; errors are reported
for every output field and for the run-level diagnostics (domain maximum of
w, RMS theta', accumulated surface rain).

  python -m nimbus_amd.accuracy [--workload c1] [--mode hevi|explicit] [--steps N]
"""
import argparse
import json
import time

import numpy as np

from . import lib as L
from . import workloads as W


def run(ctx, wl, precision, mode, steps, graphs=True):
    """One member alone for `steps` large steps: (fields, diagnostics, seconds)."""
    dtype = np.float32 if precision == L.PRECISION_MIXED else np.float64
    m = ctx.member(wl.grid, wl.config(acoustic_mode=mode), wl.base(), precision)
    try:
        m.upload(wl.initial_state(dtype))
        ctx.set_graphs(graphs)
        t0 = time.perf_counter()
        done = 0
        while done < steps:  # chunks keep each call short
            n = min(200, steps - done)
            ctx.step([m], n)
            done += n
        seconds = time.perf_counter() - t0
        return m.download(), m.diag(), seconds
    finally:
        ctx.set_graphs(False)
        m.destroy()


def compare(wl, mixed, fp64, d_mixed, d_fp64):
    """Per-field and diagnostic errors of `mixed` against `fp64` on the physical cells."""
    base = wl.base()
    phys = (slice(1, -1), slice(1, -1), slice(None))
    out = {"fields_max_abs": {}, "fields_rms": {}}
    for f in L.FIELDS:
        d = np.asarray(mixed[f], np.float64)[phys] - np.asarray(fp64[f], np.float64)[phys]
        out["fields_max_abs"][f] = float(np.max(np.abs(d)))
        out["fields_rms"][f] = float(np.sqrt(np.mean(d * d)))
    th0 = base["theta0"][None, None, :]
    tp_m = np.asarray(mixed["theta"], np.float64)[phys] - th0
    tp_d = np.asarray(fp64["theta"], np.float64)[phys] - th0
    rain_m = float(np.sum(mixed["rain_mm"][1:-1, 1:-1]))
    rain_d = float(np.sum(fp64["rain_mm"][1:-1, 1:-1]))
    out.update({
        "max_w_mixed": d_mixed["max_w"], "max_w_fp64": d_fp64["max_w"],
        "max_w_error": abs(d_mixed["max_w"] - d_fp64["max_w"]),
        "w_pointwise_max_error": out["fields_max_abs"]["w"],
        "rms_theta_p_mixed": d_mixed["rms_theta_p"], "rms_theta_p_fp64": d_fp64["rms_theta_p"],
        "rms_theta_p_error": abs(d_mixed["rms_theta_p"] - d_fp64["rms_theta_p"]),
        "theta_p_rms_difference": float(np.sqrt(np.mean((tp_m - tp_d) ** 2))),
        "rain_total_mm_mixed": rain_m, "rain_total_mm_fp64": rain_d,
        "rain_relative_error": abs(rain_m - rain_d) / max(abs(rain_d), 1e-30),
    })
    out["targets_met"] = bool(out["max_w_error"] < 0.5 and out["rms_theta_p_error"] < 0.05)
    return out


def one_hour(ctx, wl, mode=L.ACOUSTIC_HEVI, steps=None, every_seconds=600.0, control=True, progress=None):
    """Mixed and FP64 members advanced side by side from the same state; the
    run-level errors at every `every_seconds` of simulated time, and the full
    per-field comparison at the end.

    control: also advance an FP64 member from the FP32-rounded initial state
    (the state the mixed member starts from) and report its distance to the
    FP64 run beside the mixed one. It separates what FP32 arithmetic adds
    from how far the flow itself carries a perturbation the size of one FP32
    rounding of the initial fields. progress(simulated_seconds) is called
    after every comparison."""
    steps = int(steps or round(3600.0 / wl.dt))
    chunk = max(1, int(round(every_seconds / wl.dt)))
    cfg = wl.config(acoustic_mode=mode)
    runs = [("mixed", L.PRECISION_MIXED, np.float32), ("fp64", L.PRECISION_FP64, np.float64)]
    if control:
        runs.append(("control", L.PRECISION_FP64, np.float32))
    ms = {}
    series = []
    wall = {name: 0.0 for name, _, _ in runs}
    try:
        for name, prec, ic_dtype in runs:
            ms[name] = ctx.member(wl.grid, cfg, wl.base(), prec)
            ms[name].upload(wl.initial_state(ic_dtype))
        ctx.set_graphs(True)
        done = 0
        while done < steps:
            n = min(chunk, steps - done)
            for name, m in ms.items():
                t0 = time.perf_counter()
                ctx.step([m], n)
                wall[name] += time.perf_counter() - t0
            done += n
            dm, dd = ms["mixed"].diag(), ms["fp64"].diag()
            e = {"simulated_seconds": done * wl.dt, "max_w_mixed": dm["max_w"], "max_w_fp64": dd["max_w"],
                 "max_w_error": abs(dm["max_w"] - dd["max_w"]),
                 "rms_theta_p_error": abs(dm["rms_theta_p"] - dd["rms_theta_p"]),
                 "targets_met": bool(abs(dm["max_w"] - dd["max_w"]) < 0.5 and
                                     abs(dm["rms_theta_p"] - dd["rms_theta_p"]) < 0.05)}
            if control:
                dc = ms["control"].diag()
                e["control_max_w_error"] = abs(dc["max_w"] - dd["max_w"])
                e["control_rms_theta_p_error"] = abs(dc["rms_theta_p"] - dd["rms_theta_p"])
            series.append(e)
            if progress:
                progress(done * wl.dt)
        fp64 = ms["fp64"].download()
        res = compare(wl, ms["mixed"].download(), fp64, ms["mixed"].diag(), ms["fp64"].diag())
        if control:
            c = compare(wl, ms["control"].download(), fp64, ms["control"].diag(), ms["fp64"].diag())
            res["control_fp64_from_f32_state"] = {k: c[k] for k in (
                "max_w_error", "w_pointwise_max_error", "rms_theta_p_error", "theta_p_rms_difference",
                "rain_relative_error")}
    finally:
        ctx.set_graphs(False)
        for m in ms.values():
            m.destroy()
    met = [e["simulated_seconds"] for e in series if e["targets_met"]]
    first_fail = next((e["simulated_seconds"] for e in series if not e["targets_met"]), None)
    res.update({"workload": wl.describe(), "simulated_seconds": steps * wl.dt, "large_steps": steps,
                "acoustic_mode": "hevi" if mode == L.ACOUSTIC_HEVI else "explicit",
                "wall_seconds_mixed": round(wall["mixed"], 3),
                "wall_seconds_fp64": round(wall["fp64"], 3),
                "series": series, "targets_met_through_seconds": (first_fail - chunk * wl.dt) if first_fail
                else (met[-1] if met else 0.0)})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--mode", default="hevi", choices=["hevi", "explicit"])
    ap.add_argument("--steps", type=int, default=0)
    args = ap.parse_args()
    ctx = L.Context(0)
    mode = L.ACOUSTIC_HEVI if args.mode == "hevi" else L.ACOUSTIC_EXPLICIT
    print(json.dumps(one_hour(ctx, W.NAMED[args.workload], mode, args.steps or None)))
    ctx.close()


if __name__ == "__main__":
    main()
