"""Ensemble serving: continuous batching of member runs on one device.

Requests (grid, perturbed sounding, simulated duration) arrive over time. A
request is admitted as soon as it has arrived and its device footprint fits
the HBM budget next to the members already resident (first come, first
served: a request that does not fit holds back the ones behind it). All
resident members advance together as one batch, so every kernel launch of a
large step covers all of them; each round advances the batch by
min(quantum, fewest steps any member still needs) large steps, after which
finished members report their diagnostics, free their memory and make room
for waiting requests. A member's arithmetic does not depend on what else is
in the batch, so a result is bitwise the same alone and in the mix.

The initial state of an arrived request is generated on a host thread while
the device works on the current batch: a round's steps are enqueued without
waiting (nimbus_hip_step_async), arrivals are collected and their host
preparation started, and only then does the loop wait for the device.

`max_resident_cells` caps the cells resident at once besides the HBM budget
(None: the budget alone decides). A batch of a few tens of million cells
already keeps the device busy; admitting more only spreads the same
throughput over more members, so under overload every request waits longer
and each membership change rebuilds a larger batch. A request larger than
the cap is admitted when nothing else is resident.

`admission` orders the waiting requests: "fifo" (first come, first served)
or "shortest" (least work, cells x steps, first: under overload the short
requests stop queueing behind long ones, which lowers the median completion
time, and a long request can wait as long as shorter ones keep arriving).
Evicted members go back to the head of the queue under either order.

With a TuningCache, every admitted member gets the autotuned launch geometry
of its shape (tuned the first time the shape is seen); the geometry never
changes results. `column_solve` fixes the HE-VI column-solve form of every
member the server creates (None keeps each member's own default); it is part
of the member's arithmetic, so run_alone must be given the same value.

With `preempt` on, a request that does not fit may evict resident members
whose remaining work (cells x steps) is at least `preempt_ratio` times its
own: each is snapshotted to host memory (nimbus_hip_member_snapshot),
destroyed, and queued behind the request; when it is admitted again it is
restored and continues bit for bit where it stopped. A member is evicted at
most once. With `results_dir` set, every finished member's fields are
delivered as .npy files (nimbus_hip_member_export_npy) in
results_dir/req_<rid>/.
"""
import collections
import concurrent.futures
import dataclasses
import math
import os
import time

import numpy as np

from . import lib as L
from . import workloads as W


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    workload: W.Workload    # grid, dt, substeps, steps, sounding, seed/member


def poisson_requests(seed=4, rate=2.0, horizon=600.0, sizes=(48, 64, 96, 128, 160, 192, 224, 256),
                     levels=(40, 60, 80), dx_range=(500.0, 1000.0), minutes=(15.0, 90.0), substeps=6):
    """Seeded Poisson arrivals (rate per second over `horizon` seconds). Each
    request draws nx = ny from `sizes`, nz from `levels`, dx uniform in
    `dx_range` (dt = workloads.long_run_dt(dx, dz), 20 km top), a perturbed sounding and a
    simulated duration uniform in `minutes`."""
    s = W.SeededStream(seed)
    out = []
    t = 0.0
    while True:
        t += -math.log(max(1.0 - s.uniform(), 1e-300)) / rate
        if t > horizon:
            break
        n = s.choice(list(sizes))
        nz = s.choice(list(levels))
        dx = round(s.between(*dx_range), 1)
        dt = W.long_run_dt(dx, 20000.0 / nz)
        duration = 60.0 * s.between(*minutes)
        rid = len(out)
        wl = W.Workload(f"req{rid}", n, n, nz, dx, 20000.0 / nz, dt, substeps, max(1, int(round(duration / dt))),
                        seed=seed, member=rid, sounding=W.perturbed_sounding(s))
        out.append(Request(rid, t, wl))
    return out


@dataclasses.dataclass
class Result:
    rid: int
    cells: int
    steps: int
    arrival: float
    admitted: float
    completed: float
    diag: dict
    fields: dict = None

    @property
    def latency(self):
        return self.completed - self.arrival


@dataclasses.dataclass
class Report:
    results: list
    rejected: list
    wall_seconds: float
    device_step_seconds: float
    rounds: int
    peak_bytes: int
    peak_members: int
    preemptions: int = 0
    unfinished: list = dataclasses.field(default_factory=list)

    @property
    def cell_updates(self):
        return sum(r.cells * r.steps for r in self.results)

    def latency_percentile(self, q):
        lat = sorted(r.latency for r in self.results)
        if not lat:
            return float("nan")
        return float(np.percentile(np.array(lat), q))

    def summary(self):
        return {"requests": len(self.results), "rejected": len(self.rejected),
                "cell_updates": self.cell_updates, "wall_seconds": round(self.wall_seconds, 3),
                "cell_updates_per_s": self.cell_updates / self.wall_seconds if self.wall_seconds > 0 else 0.0,
                "device_step_seconds": round(self.device_step_seconds, 3),
                "p50_latency_s": round(self.latency_percentile(50), 3),
                "p99_latency_s": round(self.latency_percentile(99), 3),
                "rounds": self.rounds, "peak_resident_bytes": self.peak_bytes,
                "peak_resident_members": self.peak_members, "preemptions": self.preemptions,
                "unfinished": len(self.unfinished),
                "unfinished_resident": sum(1 for u in self.unfinished if u["state"] == "resident"),
                "unfinished_waiting": sum(1 for u in self.unfinished if u["state"] == "waiting"),
                "unfinished_oldest_age_s": max((u["age_s"] for u in self.unfinished), default=0.0),
                # finished work plus the steps already taken by members still resident
                "cell_updates_incl_partial": self.cell_updates_incl_partial,
                "cell_updates_per_s_incl_partial":
                    self.cell_updates_incl_partial / self.wall_seconds if self.wall_seconds > 0 else 0.0}

    @property
    def cell_updates_incl_partial(self):
        return self.cell_updates + sum((u["steps"] - u["steps_left"]) * u["cells"] for u in self.unfinished)


class _Slot:
    __slots__ = ("req", "member", "remaining", "bytes", "admitted", "evicted")

    def __init__(self, req, member, nbytes, admitted):
        self.req = req
        self.member = member
        self.remaining = req.workload.steps
        self.bytes = nbytes
        self.admitted = admitted
        self.evicted = False

    @property
    def work(self):
        return self.remaining * self.req.workload.cells


class _Suspended:
    """An evicted member: its host snapshot and the slot it came from."""
    __slots__ = ("slot", "blob")

    def __init__(self, slot, blob):
        self.slot = slot
        self.blob = blob

    @property
    def req(self):
        return self.slot.req


class Server:
    """Continuous-batching request loop over one device context."""

    def __init__(self, ctx, budget_bytes=200e9, precision=L.PRECISION_MIXED, acoustic_mode=L.ACOUSTIC_HEVI,
                 quantum=16, graphs=True, host_workers=4, preempt=False, preempt_ratio=4.0, results_dir=None,
                 tuning=None, column_solve=None, progress=None, max_resident_cells=None, admission="fifo"):
        if admission not in ("fifo", "shortest"):
            raise ValueError(f"admission must be 'fifo' or 'shortest', not {admission!r}")
        self.ctx = ctx
        self.admission = admission
        self.budget = int(budget_bytes)
        self.max_cells = None if max_resident_cells is None else int(max_resident_cells)
        self.precision = precision
        self.mode = acoustic_mode
        self.quantum = int(quantum)
        self.graphs = graphs
        self.host_workers = host_workers
        self.dtype = np.float32 if precision == L.PRECISION_MIXED else np.float64
        self.preempt = bool(preempt)
        self.preempt_ratio = float(preempt_ratio)
        self.results_dir = results_dir
        self.tuning = tuning
        self.column_solve = column_solve
        self.progress = progress
        self.preemptions = 0
        self._bytes = {}

    def footprint(self, req):
        wl = req.workload
        key = (wl.nx, wl.ny, wl.nz)
        if key not in self._bytes:
            self._bytes[key] = L.member_bytes(wl.grid, wl.config(acoustic_mode=self.mode), wl.base(),
                                              self.precision)
        return self._bytes[key]

    def _prepare(self, req):
        wl = req.workload
        return wl.base(), wl.initial_state(self.dtype)

    def run(self, requests, time_scale=1.0, keep_fields=(), clock=time.perf_counter, sleep=time.sleep,
            stop_after=None):
        """Serve `requests` (arrival times in seconds, compressed by `time_scale`).
        Returns a Report; results in completion order. `keep_fields` lists
        request ids whose final fields are downloaded. With `stop_after`
        (seconds of wall time) the run ends there and the Report lists the
        requests still resident, waiting or not yet arrived as `unfinished`."""
        keep = set(keep_fields)
        self.preemptions = 0
        queue = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting = collections.deque()
        active = []
        results, rejected = [], []
        used = peak_bytes = peak_members = rounds = 0
        dev_seconds = 0.0
        pool = concurrent.futures.ThreadPoolExecutor(max_workers=max(1, self.host_workers))
        prepared = {}
        self.ctx.set_graphs(self.graphs)
        start = clock()
        now = lambda: (clock() - start) * time_scale
        stopped = False
        try:
            while queue or waiting or active:
                if stop_after is not None and clock() - start >= stop_after:
                    stopped = True
                    break
                t = now()
                while queue and queue[0].arrival <= t:
                    r = queue.popleft()
                    prepared[r.rid] = pool.submit(self._prepare, r)
                    waiting.append(r)
                if self.admission == "shortest" and len(waiting) > 1:
                    # evicted members stay in front; the new requests by work
                    held = [w for w in waiting if isinstance(w, _Suspended)]
                    fresh = sorted((w for w in waiting if not isinstance(w, _Suspended)),
                                   key=lambda r: (r.workload.cells * r.workload.steps, r.arrival, r.rid))
                    waiting = collections.deque(held + fresh)
                while waiting:
                    item = waiting[0]
                    if isinstance(item, _Suspended):
                        if used + item.slot.bytes > self.budget:
                            break
                        waiting.popleft()
                        item.slot.member = self.ctx.restore(item.blob)
                        active.append(item.slot)
                        used += item.slot.bytes
                        continue
                    r = item
                    if (self.max_cells is not None and active and
                            sum(a.req.workload.cells for a in active) + r.workload.cells > self.max_cells):
                        break
                    nb = self.footprint(r)
                    if nb > self.budget:
                        rejected.append(r.rid)
                        waiting.popleft()
                        prepared.pop(r.rid).cancel()
                        continue
                    if used + nb > self.budget:
                        victims = self._victims(active, r, used + nb - self.budget)
                        if not victims:
                            break
                        for v in victims:
                            blob = v.member.snapshot()
                            v.member.destroy()
                            v.member = None
                            v.evicted = True
                            active.remove(v)
                            used -= v.bytes
                            self.preemptions += 1
                            waiting.insert(1, _Suspended(v, blob))
                    waiting.popleft()
                    base, st = prepared.pop(r.rid).result()
                    wl = r.workload
                    m = self.ctx.member(wl.grid, wl.config(acoustic_mode=self.mode), base, self.precision)
                    configure_member(m, self.mode, self.column_solve, self.tuning)
                    m.upload(st)
                    active.append(_Slot(r, m, nb, now()))
                    used += nb
                peak_bytes = max(peak_bytes, used)
                peak_members = max(peak_members, len(active))
                if not active:
                    if queue:
                        wait = (queue[0].arrival - now()) / time_scale
                        if wait > 0:
                            sleep(min(wait, 0.05))
                    continue
                n = min(self.quantum, min(s.remaining for s in active))
                self.ctx.step([s.member for s in active], n, wait=False)
                # while the device works: collect arrivals and start their host preparation
                t = now()
                while queue and queue[0].arrival <= t:
                    r = queue.popleft()
                    prepared[r.rid] = pool.submit(self._prepare, r)
                    waiting.append(r)
                self.ctx.synchronize()
                dev_seconds += self.ctx.last_step_ms() / 1e3
                rounds += 1
                if self.progress is not None and rounds % 25 == 0:
                    self.progress(f"round {rounds}: t={now():.1f} s, resident {len(active)}, waiting "
                                  f"{len(waiting)}, done {len(results)}, not arrived {len(queue)}")
                still = []
                for s in active:
                    s.remaining -= n
                    if s.remaining > 0:
                        still.append(s)
                        continue
                    d = s.member.diag()
                    fields = s.member.download() if s.req.rid in keep else None
                    if self.results_dir is not None:
                        out = os.path.join(self.results_dir, f"req_{s.req.rid}")
                        os.makedirs(out, exist_ok=True)
                        s.member.export_npy(out)
                    wl = s.req.workload
                    results.append(Result(s.req.rid, wl.cells, wl.steps, s.req.arrival, s.admitted, now(), d,
                                          fields))
                    s.member.destroy()
                    used -= s.bytes
                active = still
            unfinished = []
            if stopped:
                t = now()
                for s in active:
                    unfinished.append({"rid": s.req.rid, "state": "resident", "age_s": round(t - s.req.arrival, 3),
                                       "steps_left": s.remaining, "steps": s.req.workload.steps,
                                       "cells": s.req.workload.cells})
                for item in waiting:
                    r = item.req if isinstance(item, _Suspended) else item
                    unfinished.append({"rid": r.rid, "state": "waiting", "age_s": round(t - r.arrival, 3),
                                       "steps_left": r.workload.steps, "steps": r.workload.steps,
                                       "cells": r.workload.cells})
                for r in queue:
                    unfinished.append({"rid": r.rid, "state": "not arrived", "age_s": 0.0,
                                       "steps_left": r.workload.steps, "steps": r.workload.steps,
                                       "cells": r.workload.cells})
        finally:
            for s in active:
                s.member.destroy()
            waiting.clear()
            pool.shutdown(wait=True, cancel_futures=True)
            self.ctx.set_graphs(False)
        wall = (clock() - start)
        rep = Report(results, rejected, wall, dev_seconds, rounds, peak_bytes, peak_members, self.preemptions)
        rep.unfinished = unfinished
        return rep

    def _victims(self, active, req, need):
        """Resident members to evict so that `need` more bytes fit, largest
        remaining work first, each with at least preempt_ratio times the
        request's work and not evicted before; [] when they cannot free enough."""
        if not self.preempt:
            return []
        work = req.workload.steps * req.workload.cells
        cand = sorted((s for s in active if not s.evicted and s.work >= self.preempt_ratio * work),
                      key=lambda s: -s.work)
        out, freed = [], 0
        for s in cand:
            if freed >= need:
                break
            out.append(s)
            freed += s.bytes
        return out if freed >= need else []


def configure_member(m, mode, column_solve=None, tuning=None):
    """The server's per-member settings: the HE-VI column-solve form (part of
    the member's arithmetic) and the cached launch geometry (not)."""
    if column_solve is not None and mode == L.ACOUSTIC_HEVI:
        m.set_column_solve(column_solve)
    if tuning is not None:
        tuning.apply(m, mode)


def run_alone(ctx, req, precision=L.PRECISION_MIXED, acoustic_mode=L.ACOUSTIC_HEVI, graphs=True, column_solve=None,
              tuning=None):
    """One request on its own (the reference for 'same alone and in the mix' and FP64 reruns)."""
    wl = req.workload
    dtype = np.float32 if precision == L.PRECISION_MIXED else np.float64
    m = ctx.member(wl.grid, wl.config(acoustic_mode=acoustic_mode), wl.base(), precision)
    # the column-solve form applies to the FP32-storage members it was chosen
    # for; FP64 reruns keep their default (partitioned up to 129 levels)
    configure_member(m, acoustic_mode, column_solve if precision == L.PRECISION_MIXED else None, tuning)
    ctx.set_graphs(graphs)
    try:
        m.upload(wl.initial_state(dtype))
        ctx.step([m], wl.steps)
        return m.download(), m.diag()
    finally:
        ctx.set_graphs(False)
        m.destroy()
