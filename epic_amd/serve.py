"""Served ensemble: a request loop with continuous batching.

Requests are member runs (a grid, its parcels, a fixed dt and a number of
RK steps) that arrive over time.  The server keeps every admitted member's
parcel and field state resident in HBM and advances all of them together
with one batched LS-RK step per tick (``core.Batch``: one launch per kernel
over all members, one captured HIP graph).  Between ticks it

* retires the members that have taken their requested steps (their final
  parcel state becomes the result, their device memory is freed),
* admits waiting requests in arrival order while the HBM budget
  (``core.member_bytes`` of every resident member) and the slot limit allow,

so requests arriving while others are on the GPU join the running batch as
soon as memory frees up.  A member's arithmetic does not depend on the other
members of its batch, so every result equals the member's run alone.

Reported: per-tick step latency (p50/p99), per-request queueing and
completion times, and the aggregate parcel-steps/s.
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque

import numpy as np

from .core import Batch, Member, member_bytes


@dataclasses.dataclass
class Request:
    rid: int
    workload: object          # epic_amd.workloads.Member
    steps: int
    dt: float = 0.01
    arrival: float = 0.0      # seconds after the server starts
    rk_order: int = 4


@dataclasses.dataclass
class Result:
    rid: int
    parcels: dict
    t: float
    admitted: float
    finished: float
    arrival: float


class Server:
    def __init__(self, hbm_budget_bytes: int, max_members: int = 64, autotune: bool = False):
        self.budget = int(hbm_budget_bytes)
        self.max_members = int(max_members)
        self.autotune = autotune
        self._tuned = {}

    def _bytes(self, req):
        w = req.workload
        return member_bytes(w.nx, w.ny, w.nz, w.n)

    def run(self, requests, clock=time.perf_counter, idle_sleep=1e-3):
        """Serves ``requests`` (any order; admitted by arrival time) to
        completion.  Returns (results by rid, stats)."""
        waiting = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        active = []            # [req, member, t, steps_done, admitted_at]
        used = 0
        results = {}
        tick_ms, tick_parcels, tick_members = [], [], []
        batch, batch_key = None, None
        t0 = clock()
        while waiting or active:
            now = clock() - t0
            # admission: arrival order, HBM budget and slot limit
            while (waiting and waiting[0].arrival <= now and len(active) < self.max_members
                   and used + self._bytes(waiting[0]) <= self.budget):
                req = waiting.popleft()
                m = Member.from_workload(req.workload, rk_order=req.rk_order)
                if self.autotune:
                    key = (req.workload.nx, req.workload.ny, req.workload.nz, req.workload.n)
                    if key not in self._tuned:
                        self._tuned[key] = m.autotune()
                    m.set_tiling(*self._tuned[key])
                used += self._bytes(req)
                active.append([req, m, 0.0, 0, now])
            if waiting and not active and self._bytes(waiting[0]) > self.budget:
                raise ValueError(f"request {waiting[0].rid} does not fit the HBM budget")
            if not active:
                time.sleep(idle_sleep)
                continue
            key = tuple(id(a[1]) for a in active)
            if key != batch_key:
                if batch is not None:
                    batch.close()
                batch = Batch([a[1] for a in active])
                batch_key = key
            ts = np.array([a[2] for a in active])
            dts = np.array([a[0].dt for a in active])
            a0 = clock()
            ts, status = batch.step(ts, dts)
            tick_ms.append((clock() - a0) * 1e3)
            tick_parcels.append(sum(a[0].workload.n for a in active))
            tick_members.append(len(active))
            done = []
            for a, t in zip(active, ts):
                a[2] = float(t)
                a[3] += 1
                if a[3] >= a[0].steps:
                    done.append(a)
            for a in done:
                req, m = a[0], a[1]
                results[req.rid] = Result(req.rid, m.get_parcels(), a[2], a[4], clock() - t0, req.arrival)
                active.remove(a)
                used -= self._bytes(req)
            if done:
                if batch is not None:
                    batch.close()
                batch, batch_key = None, None
                for a in done:
                    a[1].close()
        wall = clock() - t0
        if batch is not None:
            batch.close()
        lat = np.array(tick_ms) if tick_ms else np.zeros(1)
        parcel_steps = float(np.sum(tick_parcels))
        busy = float(np.sum(tick_ms)) * 1e-3
        stats = {
            "ticks": len(tick_ms),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "parcel_steps": parcel_steps,
            "wall_s": wall,
            "parcel_steps_per_s_wall": parcel_steps / wall if wall > 0 else 0.0,
            "parcel_steps_per_s_busy": parcel_steps / busy if busy > 0 else 0.0,
            "mean_members_per_tick": float(np.mean(tick_members)) if tick_members else 0.0,
            "queue_wait_s_p50": float(np.percentile([r.admitted - r.arrival for r in results.values()], 50))
            if results else 0.0,
        }
        return results, stats


def poisson_requests(workloads, rate, steps, dt=0.01, seed=0):
    """Requests for the given workloads with exponential inter-arrival times
    of mean 1/rate seconds; ``steps`` is an int or a (lo, hi) range."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i, w in enumerate(workloads):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        k = steps if isinstance(steps, int) else int(rng.integers(steps[0], steps[1] + 1))
        out.append(Request(i, w, k, dt, t))
    return out
