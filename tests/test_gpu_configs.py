"""The benchmark configurations on the device against the oracle, at their
full sizes, in the precision and method each names:

- 64 x 64 x 40, dx = 1 km: the plain method (FP64, explicit acoustic
  substeps) for its 10 large steps, with the reference's substep promotion
  (N = 10);
- 160 x 160 x 80 (500 m) and 400 x 400 x 100 (250 m): two FP64 large steps in
  both acoustic modes, and the mixed path against the FP64 oracle at its FP32
  rounding bound (at 400 x 400 the FP32 member takes the sequential column
  sweep);
- the 32-member uneven ensemble in one batch, each member bitwise equal to
  itself alone, the two smallest against the oracle;
- the serving stream's request shapes (seeded Poisson arrivals, up to
  256 x 256 x 80) served together, each equal to itself alone.

The oracle runs on the host's cores (OpenMP) in the same process.
"""
import numpy as np
import pytest

from nimbus_amd import lib as L
from nimbus_amd import serving as SV
from nimbus_amd import workloads as W
from oracle import oracle_py as O
from tests.helpers import max_abs, oracle_grid
from tests.test_gpu_parity import check_fp64, check_mixed

pytestmark = pytest.mark.gpu


def _run(ctx, wl, cfg, precision, n_steps, column_solve=None):
    dtype = np.float32 if precision == L.PRECISION_MIXED else np.float64
    m = ctx.member(wl.grid, cfg, wl.base(), precision)
    try:
        if column_solve is not None:
            m.set_column_solve(column_solve)
        m.upload(wl.initial_state(dtype))
        ctx.step([m], n_steps)
        return m.download(), m.diag(), m.substeps
    finally:
        m.destroy()


def test_c0_plain_fp64_method_ten_steps_matches_oracle(hip_ctx):
    wl = W.NAMED["c0"]
    cfg = wl.config(acoustic_mode=L.ACOUSTIC_EXPLICIT)
    print(f"case: c0 {wl.describe()} fp64 explicit, {wl.steps} steps", flush=True)
    ref, n_sub = O.large_steps(oracle_grid(wl.grid), wl.base(), O.from_step_config(cfg),
                               wl.initial_state(np.float64), wl.steps, np.float64)
    out, diag, n_dev = _run(hip_ctx, wl, cfg, L.PRECISION_FP64, wl.steps)
    assert n_sub == 10 and n_dev == 10  # 6 requested, promoted by the acoustic CFL (c dt / (0.5 dz) + 1)
    check_fp64(out, ref, "c0 fp64 explicit", True, diffusive=True)
    assert max_abs(out["rain_mm"], ref["rain_mm"]) <= 1e-9 * max(1.0, float(np.max(ref["rain_mm"])))
    assert diag["steps"] == wl.steps and diag["max_w"] > 0.1


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_single_member_configs_fp64_and_mixed_match_oracle(hip_ctx, name):
    wl = W.NAMED[name]
    n_steps = 2
    og = oracle_grid(wl.grid)
    for mode in (L.ACOUSTIC_EXPLICIT, L.ACOUSTIC_HEVI):
        cfg = wl.config(acoustic_mode=mode)
        print(f"case: {name} {wl.describe()} mode {mode}, {n_steps} steps", flush=True)
        ref, _ = O.large_steps(og, wl.base(), O.from_step_config(cfg), wl.initial_state(np.float64), n_steps,
                               np.float64)
        out, _, _ = _run(hip_ctx, wl, cfg, L.PRECISION_FP64, n_steps)
        check_fp64(out, ref, f"{name} fp64 mode {mode}", True, diffusive=True)
        if mode == L.ACOUSTIC_HEVI:
            outm, _, _ = _run(hip_ctx, wl, cfg, L.PRECISION_MIXED, n_steps)
            check_mixed(outm, ref, f"{name} mixed hevi")
    if name == "c2":
        # the FP32 member of this size takes the sequential column sweep by default
        m = hip_ctx.member(wl.grid, wl.config(acoustic_mode=L.ACOUSTIC_HEVI), wl.base(), L.PRECISION_MIXED)
        try:
            assert m.column_solve() == L.COLUMN_SOLVE_SEQUENTIAL
        finally:
            m.destroy()


def test_c3_ensemble_in_one_batch_equals_members_alone(hip_ctx):
    members = W.ensemble_c3()
    n_steps = 3
    print(f"case: c3 ensemble of {len(members)} members, {n_steps} steps", flush=True)
    cfgs = [wl.config(acoustic_mode=L.ACOUSTIC_HEVI) for wl in members]
    batch = [hip_ctx.member(wl.grid, c, wl.base(), L.PRECISION_MIXED) for wl, c in zip(members, cfgs)]
    try:
        for m, wl in zip(batch, members):
            m.upload(wl.initial_state(np.float32))
        hip_ctx.set_graphs(True)
        hip_ctx.step(batch, n_steps)
        together = [m.download() for m in batch]
    finally:
        hip_ctx.set_graphs(False)
        for m in batch:
            m.destroy()
    for q, (wl, c) in enumerate(zip(members, cfgs)):
        alone, _, _ = _run(hip_ctx, wl, c, L.PRECISION_MIXED, n_steps)
        for f in list(L.FIELDS) + ["rain_mm"]:
            assert np.array_equal(alone[f], together[q][f]), (q, f)
    smallest = sorted(range(len(members)), key=lambda q: members[q].cells)[:2]
    for q in smallest:
        wl = members[q]
        ref, _ = O.large_steps(oracle_grid(wl.grid), wl.base(), O.from_step_config(cfgs[q]),
                               wl.initial_state(np.float64), n_steps, np.float64)
        check_mixed(together[q], ref, f"c3 member {q} {wl.describe()}")


def test_c4_request_shapes_served_together_equal_alone(hip_ctx):
    stream = SV.poisson_requests(seed=4, rate=2.0, horizon=600.0)
    picked = stream[:10]
    largest = next(r for r in stream if r.workload.nx == 256 and r.workload.nz == 80)
    if largest not in picked:
        picked.append(largest)
    reqs = []
    for q, r in enumerate(picked):
        wl = r.workload
        wl = W.Workload(wl.name, wl.nx, wl.ny, wl.nz, wl.dx, wl.dz, wl.dt, wl.substeps, 3 + q % 3, seed=wl.seed,
                        member=wl.member, sounding=wl.sounding)
        reqs.append(SV.Request(q, 0.0, wl))
    print("case: c4 shapes " + ", ".join(r.workload.describe() for r in reqs), flush=True)
    srv = SV.Server(hip_ctx, quantum=2, host_workers=4)
    rep = srv.run(reqs, keep_fields=[r.rid for r in reqs])
    assert len(rep.results) == len(reqs) and rep.peak_members >= 2
    by_id = {r.rid: r for r in reqs}
    for res in rep.results:
        fields, diag = SV.run_alone(hip_ctx, by_id[res.rid])
        for f in list(L.FIELDS) + ["rain_mm"]:
            assert np.array_equal(res.fields[f], fields[f]), (res.rid, f)
        assert res.diag["max_w"] == diag["max_w"]


def test_state_dependent_microphysics_subcycling_matches_oracle(hip_ctx):
    """A large step longer than the sedimentation limit of its state
    (0.8 dz / speed of the cell's total water): the device clock takes the
    reference's steps min(remaining, limit), the limit recomputed after each,
    and the FP64 member equals the FP64 oracle, which follows the same rule."""
    from tests.helpers import small_case
    g, base, st = small_case(nx=14, ny=12, nz=20, dz=100.0, cloudy=True, dtype=np.float64)
    cfg = L.step_config(12.0, 6, acoustic_mode=L.ACOUSTIC_HEVI, lateral=L.BC_RIGID_SLIP)
    og = oracle_grid(g)
    # the state's limit is below dt: the interval is subcycled
    _, lengths = O.microphysics_step(og, O.from_step_config(cfg), st, cfg.dt, np.float64, lengths=True)
    print(f"case: micro subcycling, oracle lengths on the initial state {list(lengths)}", flush=True)
    assert len(lengths) >= 2 and len(set(np.round(lengths, 12))) >= 2
    n_steps = 2
    ref, _ = O.large_steps(og, base, O.from_step_config(cfg), st, n_steps, np.float64)
    m = hip_ctx.member(g, cfg, base, L.PRECISION_FP64)
    try:
        m.upload(st)
        hip_ctx.set_graphs(True)
        hip_ctx.step([m], n_steps)
        out = m.download()
    finally:
        hip_ctx.set_graphs(False)
        m.destroy()
    check_fp64(out, ref, "micro subcycling", False)
    assert max_abs(out["rain_mm"], ref["rain_mm"]) <= 1e-9 * max(1.0, float(np.max(ref["rain_mm"])))
