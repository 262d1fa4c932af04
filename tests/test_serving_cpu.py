"""Serving loop logic on the CPU with a stand-in device context: admission by
HBM budget in arrival order, rejection of requests larger than the budget,
step accounting per member and latency percentiles. No device is touched."""
import os

import numpy as np
import pytest

from nimbus_amd import lib as L
from nimbus_amd import serving as SV
from nimbus_amd import workloads as W


class FakeMember:
    def __init__(self, ctx, grid, config, base, precision):
        self.ctx = ctx
        self.grid = grid
        self.steps = 0
        self.destroyed = False
        self.uploaded = False
        ctx.created.append(self)

    def upload(self, st):
        self.uploaded = True

    def download(self):
        return {"steps": self.steps}

    def diag(self):
        return {"steps": self.steps, "max_w": 0.0}

    def destroy(self):
        assert not self.destroyed
        self.destroyed = True

    def snapshot(self):
        return {"grid": self.grid, "steps": self.steps}

    def export_npy(self, directory):
        self.ctx.exported.append(directory)


class FakeContext:
    """Records batches; each step 'takes' 1 ms of device time."""

    def __init__(self):
        self.created = []
        self.restored = []
        self.exported = []
        self.batches = []
        self.ms = 0.0
        self.graphs = False
        self.pending = False

    def member(self, grid, config, base, precision=L.PRECISION_MIXED):
        return FakeMember(self, grid, config, base, precision)

    def restore(self, blob):
        m = FakeMember(self, blob["grid"], None, None, None)
        m.steps = blob["steps"]
        m.uploaded = True
        self.restored.append(m)
        return m

    def step(self, members, n_steps=1, wait=True):
        assert all(m.uploaded and not m.destroyed for m in members)
        assert not self.pending, "a step was enqueued while another was not waited for"
        self.batches.append((len(members), n_steps))
        for m in members:
            m.steps += n_steps
        self.ms = float(n_steps)
        self.pending = not wait

    def synchronize(self):
        self.pending = False

    def last_step_ms(self):
        return self.ms

    def set_graphs(self, enable):
        self.graphs = bool(enable)


def small_requests(n, steps, arrival_gap=0.0, size=48, nz=40):
    out = []
    for r in range(n):
        wl = W.Workload(f"t{r}", size, size, nz, 1000.0, 20000.0 / nz, 6.0, 6, steps[r % len(steps)],
                        seed=9, member=r)
        out.append(SV.Request(r, r * arrival_gap, wl))
    return out


def test_poisson_stream_is_seeded_and_in_range():
    a = SV.poisson_requests(seed=4, rate=2.0, horizon=60.0)
    b = SV.poisson_requests(seed=4, rate=2.0, horizon=60.0)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert 80 <= len(a) <= 160  # 120 expected
    assert all(0.0 < r.arrival <= 60.0 for r in a)
    assert all(np.diff([r.arrival for r in a]) > 0)
    for r in a:
        wl = r.workload
        assert 48 <= wl.nx <= 256 and wl.nx == wl.ny and wl.nz in (40, 60, 80)
        assert 500.0 <= wl.dx <= 1000.0
        assert 15 * 60 - wl.dt <= wl.steps * wl.dt <= 90 * 60 + wl.dt


def test_admission_respects_budget_and_order():
    reqs = small_requests(6, steps=[5, 9, 3])
    ctx = FakeContext()
    srv = SV.Server(ctx, quantum=4, host_workers=2)
    one = srv.footprint(reqs[0])
    srv.budget = 2 * one
    rep = srv.run(reqs, time_scale=1e6, sleep=lambda s: None)
    assert sorted(r.rid for r in rep.results) == list(range(6))
    assert rep.rejected == []
    assert rep.peak_members <= 2 and rep.peak_bytes <= 2 * one
    assert max(n for n, _ in ctx.batches) == 2
    # admission in arrival order: member objects were created in request order
    assert [m.grid.nx for m in ctx.created] == [48] * 6
    for res in rep.results:
        assert res.diag["steps"] == res.steps
        assert res.admitted >= res.arrival and res.completed >= res.admitted
    # every round advanced by at most the quantum and never past a member's end
    assert all(0 < n <= 4 for _, n in ctx.batches)
    assert all(m.destroyed for m in ctx.created)


def test_oversized_request_is_rejected():
    reqs = small_requests(3, steps=[2])
    big = W.Workload("big", 128, 128, 80, 1000.0, 250.0, 6.0, 6, 2, seed=1, member=7)
    reqs.append(SV.Request(3, 0.0, big))
    ctx = FakeContext()
    srv = SV.Server(ctx, quantum=8, host_workers=1)
    srv.budget = srv.footprint(reqs[0]) * 3
    rep = srv.run(reqs, time_scale=1e6, sleep=lambda s: None)
    assert rep.rejected == [3]
    assert sorted(r.rid for r in rep.results) == [0, 1, 2]


def test_latency_percentiles_and_throughput_accounting():
    reqs = small_requests(10, steps=[3, 7], arrival_gap=0.001)
    ctx = FakeContext()
    srv = SV.Server(ctx, quantum=2, host_workers=2)
    rep = srv.run(reqs, time_scale=1.0)
    s = rep.summary()
    assert s["requests"] == 10
    assert 0.0 <= s["p50_latency_s"] <= s["p99_latency_s"]
    assert s["cell_updates"] == sum(r.workload.cells * r.workload.steps for r in reqs)
    assert rep.device_step_seconds == pytest.approx(sum(n for _, n in ctx.batches) / 1e3)


def test_preemption_lets_a_short_request_through_and_resumes_the_long_one(tmp_path):
    """A long request fills the budget; a short one arrives later. Without
    preemption it waits for the long one; with it the long member is
    snapshotted, the short one runs, and the long one resumes from its
    snapshot and still takes exactly its own number of steps."""
    long_r = W.Workload("long", 48, 48, 40, 1000.0, 500.0, 6.0, 6, 400, seed=9, member=0)
    short_r = W.Workload("short", 48, 48, 40, 1000.0, 500.0, 6.0, 6, 10, seed=9, member=1)
    reqs = [SV.Request(0, 0.0, long_r), SV.Request(1, 0.0, short_r)]
    one = L.member_bytes(long_r.grid, long_r.config(acoustic_mode=L.ACOUSTIC_HEVI), long_r.base())
    for preempt in (False, True):
        ctx = FakeContext()
        srv = SV.Server(ctx, budget_bytes=int(1.5 * one), quantum=8, preempt=preempt,
                        results_dir=str(tmp_path / str(preempt)))
        # the short request is admitted only after the long one is resident
        reqs[1].arrival = 0.0
        rep = srv.run(reqs, clock=_stepping_clock(), sleep=lambda s: None)
        order = [r.rid for r in rep.results]
        by = {r.rid: r for r in rep.results}
        assert by[0].diag["steps"] == 400 and by[1].diag["steps"] == 10
        if preempt:
            assert order == [1, 0] and rep.preemptions == 1 and len(ctx.restored) == 1
        else:
            assert order == [0, 1] and rep.preemptions == 0 and not ctx.restored
        assert sorted(os.path.basename(d) for d in ctx.exported) == ["req_0", "req_1"]


def _stepping_clock():
    t = [0.0]

    def clock():
        t[0] += 1e-3
        return t[0]
    return clock


class TunableMember:
    def __init__(self, grid, precision):
        self.grid = grid
        self.precision = precision
        self.autotuned = 0
        self.tiling = None

    def autotune(self, trials):
        self.autotuned += 1
        self.tiling = (32, 3, 0)
        return 32, 3, 0, 1.25

    def set_tiling(self, w, c, r):
        self.tiling = (w, c, r)


def test_tuning_cache_tunes_each_shape_once_and_persists(tmp_path):
    from nimbus_amd import tuning as T
    cache = T.TuningCache("gfx950", directory=str(tmp_path))
    g = L.make_grid(48, 48, 40, 1000.0, 1000.0, 500.0)
    a, b = TunableMember(g, L.PRECISION_MIXED), TunableMember(g, L.PRECISION_MIXED)
    assert cache.apply(a, L.ACOUSTIC_HEVI) == (32, 3, 0) and a.autotuned == 1
    assert cache.apply(b, L.ACOUSTIC_HEVI) == (32, 3, 0) and b.autotuned == 0 and b.tiling == (32, 3, 0)
    other = TunableMember(g, L.PRECISION_FP64)
    cache.apply(other, L.ACOUSTIC_HEVI)
    assert other.autotuned == 1  # precision is part of the key
    again = T.TuningCache("gfx950", directory=str(tmp_path))
    c = TunableMember(g, L.PRECISION_MIXED)
    again.apply(c, L.ACOUSTIC_HEVI)
    assert c.autotuned == 0 and c.tiling == (32, 3, 0) and again.hits == 1


def test_stop_after_reports_the_backlog_and_the_partial_work():
    """A run stopped by wall time lists every request not finished (resident,
    waiting, not arrived) and counts the steps resident members already took."""
    reqs = small_requests(4, steps=[400])
    late = small_requests(1, steps=[5])[0]
    reqs.append(SV.Request(4, 1e9, late.workload))  # arrives long after the stop
    ctx = FakeContext()
    srv = SV.Server(ctx, quantum=4, host_workers=1)
    srv.budget = 2 * srv.footprint(reqs[0])
    ticks = iter(range(10 ** 6))
    rep = srv.run(reqs, clock=lambda: float(next(ticks)), sleep=lambda s: None, stop_after=40.0)
    s = rep.summary()
    states = sorted(u["state"] for u in rep.unfinished)
    assert s["requests"] == 0 and s["unfinished"] == 5
    assert states == ["not arrived", "resident", "resident", "waiting", "waiting"]
    taken = sum(m.steps for m in ctx.created)
    assert 0 < taken < 2 * 400
    cells = reqs[0].workload.cells
    assert s["cell_updates_incl_partial"] == taken * cells
    assert s["cell_updates_per_s_incl_partial"] == pytest.approx(taken * cells / rep.wall_seconds)
    assert all(m.destroyed for m in ctx.created)


def test_resident_cell_cap_holds_back_requests_until_cells_free():
    """With max_resident_cells the batch never holds more cells than the cap
    (one request alone may exceed it); every request still finishes, in
    arrival order of admission."""
    reqs = small_requests(6, steps=[6, 10])
    cells = reqs[0].workload.cells
    ctx = FakeContext()
    srv = SV.Server(ctx, quantum=4, host_workers=1, max_resident_cells=2 * cells + 1)
    rep = srv.run(reqs, time_scale=1e6, sleep=lambda s: None)
    assert sorted(r.rid for r in rep.results) == list(range(6))
    assert rep.peak_members == 2
    assert max(n for n, _ in ctx.batches) == 2
    admitted = sorted(rep.results, key=lambda r: r.admitted)
    assert [r.rid for r in admitted] == sorted(r.rid for r in admitted)
    # a cap below one request's size still admits it when nothing is resident
    ctx = FakeContext()
    rep = SV.Server(ctx, quantum=4, host_workers=1, max_resident_cells=cells // 2).run(
        reqs[:2], time_scale=1e6, sleep=lambda s: None)
    assert len(rep.results) == 2 and rep.peak_members == 1


def test_shortest_first_admission_orders_waiting_requests_by_work():
    """With room for one member, requests that wait are admitted least work
    first under 'shortest' and in arrival order under 'fifo'; every request
    still takes exactly its own number of steps."""
    steps = [40, 30, 5, 20, 10]
    reqs = small_requests(5, steps=steps)
    orders = {}
    for admission in ("fifo", "shortest"):
        ctx = FakeContext()
        srv = SV.Server(ctx, quantum=64, host_workers=1, admission=admission)
        srv.budget = srv.footprint(reqs[0])
        rep = srv.run(reqs, clock=_stepping_clock(), sleep=lambda s: None)
        assert all(r.diag["steps"] == r.steps for r in rep.results)
        orders[admission] = [r.rid for r in rep.results]
    # all five arrive at once and run one at a time
    assert orders["fifo"] == [0, 1, 2, 3, 4]
    assert orders["shortest"] == [2, 4, 3, 1, 0]
    with pytest.raises(ValueError):
        SV.Server(FakeContext(), admission="lifo")
