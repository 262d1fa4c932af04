"""Serving on the device: requests that arrive while others run are admitted
as memory frees up, share every launch with the resident members, and each
result is bitwise the same as the request run alone."""
import numpy as np
import pytest

from nimbus_amd import lib as L
from nimbus_amd import serving as SV


@pytest.mark.gpu
@pytest.mark.parametrize("admission", ["fifo", "shortest"])
def test_served_results_equal_alone(hip_ctx, admission):
    print(f"case: served results equal alone, admission {admission}", flush=True)
    reqs = SV.poisson_requests(seed=11, rate=40.0, horizon=0.25, sizes=(48, 64), levels=(40, 60),
                               minutes=(0.5, 1.5))
    assert len(reqs) >= 5
    srv = SV.Server(hip_ctx, quantum=3, host_workers=2, admission=admission)
    srv.budget = 3 * srv.footprint(reqs[0])  # forces waiting and admission as members finish
    rep = srv.run(reqs, time_scale=1.0, keep_fields=[r.rid for r in reqs])
    s = rep.summary()
    print("serving summary", s)
    assert s["requests"] == len(reqs) and s["rejected"] == 0
    assert rep.peak_members >= 2
    by_id = {r.rid: r for r in reqs}
    for res in rep.results:
        fields, diag = SV.run_alone(hip_ctx, by_id[res.rid])
        for f in L.FIELDS:
            assert np.array_equal(res.fields[f], fields[f]), (res.rid, f)
        assert np.array_equal(res.fields["rain_mm"], fields["rain_mm"]), res.rid
        assert res.diag["max_w"] == diag["max_w"] and res.diag["rms_theta_p"] == diag["rms_theta_p"]


@pytest.mark.gpu
def test_preempted_members_resume_bitwise(hip_ctx):
    """Long requests fill the budget, short ones arrive behind them: with
    preemption the long members are evicted to host snapshots mid-run and
    restored later, and every result still equals the request run alone."""
    from nimbus_amd import workloads as W
    reqs = []
    # all arrive at once: the long ones are admitted first, the short ones find the budget full
    for q, (n, steps, t) in enumerate([(48, 40, 0.0), (64, 36, 0.0), (48, 4, 0.0), (48, 3, 0.0)]):
        wl = W.Workload(f"p{q}", n, n, 40, 1000.0, 500.0, 6.0, 6, steps, seed=5, member=q)
        reqs.append(SV.Request(q, t, wl))
    srv = SV.Server(hip_ctx, quantum=4, host_workers=2, preempt=True)
    srv.budget = srv.footprint(reqs[0]) + srv.footprint(reqs[1])
    rep = srv.run(reqs, time_scale=1.0, keep_fields=[r.rid for r in reqs])
    print("preemption summary", rep.summary(), flush=True)
    assert len(rep.results) == 4 and rep.preemptions >= 1
    by_id = {r.rid: r for r in reqs}
    for res in rep.results:
        fields, diag = SV.run_alone(hip_ctx, by_id[res.rid])
        for f in L.FIELDS:
            assert np.array_equal(res.fields[f], fields[f]), (res.rid, f)
        assert np.array_equal(res.fields["rain_mm"], fields["rain_mm"]), res.rid
        assert res.diag["steps"] == by_id[res.rid].workload.steps


@pytest.mark.gpu
def test_served_results_delivered_as_npy(hip_ctx, tmp_path):
    """With results_dir set, every finished request's fields arrive as the
    reference's NPY field files, equal to the fields downloaded at completion."""
    from nimbus_amd import workloads as W
    reqs = [SV.Request(q, 0.0, W.Workload(f"n{q}", 48, 48, 40, 1000.0, 500.0, 3.0, 6, 3 + q, seed=6, member=q))
            for q in range(3)]
    srv = SV.Server(hip_ctx, quantum=2, host_workers=2, results_dir=str(tmp_path))
    rep = srv.run(reqs, keep_fields=[r.rid for r in reqs])
    assert len(rep.results) == 3
    for res in rep.results:
        d = tmp_path / f"req_{res.rid}"
        for f in L.FIELDS:
            a = np.load(str(d / f"{f}.npy"))
            assert a.dtype == np.float32 and np.array_equal(a, res.fields[f].astype(np.float32)), (res.rid, f)
        r = np.load(str(d / "accumulated_rainfall.npy"))
        assert np.array_equal(r, res.fields["rain_mm"].astype(np.float32))


@pytest.mark.gpu
def test_uneven_ensemble_batch_matches_members_alone(hip_ctx):
    """Members of different grids, dx and hence dt share one batch for a few steps."""
    from nimbus_amd import workloads as W
    members = W.ensemble_c3(n_members=5, seed=3)
    for wl in members:
        wl.steps = 4
    reqs = [SV.Request(q, 0.0, wl) for q, wl in enumerate(members)]
    srv = SV.Server(hip_ctx, quantum=4, host_workers=2)
    rep = srv.run(reqs, keep_fields=list(range(len(reqs))))
    assert rep.rounds == 1 and rep.peak_members == len(reqs)
    for res in rep.results:
        fields, _ = SV.run_alone(hip_ctx, reqs[res.rid])
        for f in L.FIELDS:
            assert np.array_equal(res.fields[f], fields[f]), (res.rid, f)
