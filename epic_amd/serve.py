"""Served ensemble: a request loop with continuous batching.

Requests are member runs (a grid, its parcels, a fixed dt and a number of
RK steps) that arrive over time.  The server keeps every admitted member's
parcel and field state resident in HBM.  Resident members are grouped into
cohorts; a cohort is one ``core.Batch`` (one launch per kernel over all its
members, one captured HIP graph) on its own stream.  Every tick launches the
step of every cohort without waiting (``Batch.launch``), so cohorts run
concurrently on the GPU, and while they run the server

* admits waiting requests in arrival order while the HBM budget
  (``core.member_bytes`` of every resident member) and the slot limit allow:
  it allocates the member, uploads its parcels on the member's own stream
  and captures the new cohort's step graph, all overlapping the tick (a
  retired member's memory counts against the budget until it is released);
* copies out the results of members retired at the previous tick.

Then it waits for the tick (``Batch.wait``).  With ``parcel_ops_every = K``
every member runs the parcel operations of EPIC's run loop after each K-th
of its steps, merge then split (epic3d.f90:143-148), on its own stream and
its own host thread, concurrently with the other members'; members are
allocated with room for the children of the splits.  The step graphs read
the parcel counts from device memory, so a cohort keeps its captured graph
when its members' counts change.  Long runs need the operations: without
splitting, ellipsoids stretch until their support points leave the z halo
layers.  A member that
has taken its requested steps, or whose step or parcel operation failed
(its status is returned in the result), leaves its cohort; only that
cohort's graph is rebuilt.  Device
memory is released after the wait, when no step is in flight.  A member's
arithmetic does not depend on the other members of its cohort or on the
cohorts running beside it, so every result equals the member's run alone.

Reported: per-tick step latency (p50/p99), per-request queueing and
completion times, and the aggregate parcel-steps/s (over the busy ticks and
over the wall clock).
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .core import Batch, EpicError, Member, member_bytes


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
    parcels: dict             # final parcel state (None when the run failed)
    t: float
    admitted: float
    finished: float
    arrival: float
    steps_done: int = 0
    status: int = 0           # 0, or the EPIC_E_* code of the failed step
    message: str = ""


@dataclasses.dataclass
class _Entry:
    req: Request
    member: Member
    nbytes: int
    admitted: float
    t: float = 0.0
    steps: int = 0
    status: int = 0
    message: str = ""


def _parcel_ops(e):
    """EPIC's run-loop parcel operations after a step (epic3d.f90:143-148):
    merge, then split; returns the EpicError of a failure, or None."""
    try:
        e.member.parcel_merge()
        e.member.parcel_split()
    except EpicError as err:
        return err
    return None


class _Cohort:
    def __init__(self, entries):
        self.entries = list(entries)
        self.batch = Batch([e.member for e in self.entries])
        self.batch.prepare()

    def close(self):
        self.batch.close()


class Server:
    def __init__(self, hbm_budget_bytes: int, max_members: int = 64, autotune: bool = False,
                 parcel_ops_every: int = 0, capacity_factor: float = 2.0, ops_threads: int = 8):
        self.ops_threads = max(1, int(ops_threads))
        self.budget = int(hbm_budget_bytes)
        self.max_members = int(max_members)
        self.autotune = autotune
        self.parcel_ops_every = int(parcel_ops_every)
        self.capacity_factor = float(capacity_factor) if parcel_ops_every else 1.0
        self._tuned = {}

    def _capacity(self, req):
        return int(req.workload.n * self.capacity_factor)

    def _bytes(self, req):
        w = req.workload
        return member_bytes(w.nx, w.ny, w.nz, self._capacity(req))

    def run(self, requests, clock=time.perf_counter, idle_sleep=1e-3):
        """Serves ``requests`` (any order; admitted by arrival time) to
        completion.  Returns (results by rid, stats)."""
        waiting = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        for r in waiting:
            if self._bytes(r) > self.budget:
                raise ValueError(f"request {r.rid} does not fit the HBM budget")
        cohorts = []           # list of _Cohort
        fresh = []             # admitted entries that join at the next tick
        to_collect = []        # retired entries whose results are copied out during the next tick
        to_free = []           # members released after the next wait
        used = 0
        results = {}
        tick_ms, tick_parcels, tick_members, ops_ms = [], [], [], []
        pool = ThreadPoolExecutor(max_workers=self.ops_threads)
        t0 = clock()

        def n_active():
            return sum(len(c.entries) for c in cohorts) + len(fresh)

        def admit(now):
            nonlocal used
            while (waiting and waiting[0].arrival <= now and n_active() < self.max_members
                   and used + self._bytes(waiting[0]) <= self.budget):
                req = waiting.popleft()
                m = Member.from_workload(req.workload, rk_order=req.rk_order, max_parcels=self._capacity(req))
                if self.autotune:
                    key = (req.workload.nx, req.workload.ny, req.workload.nz, req.workload.n)
                    if key not in self._tuned:
                        self._tuned[key] = m.autotune()
                    m.set_tiling(*self._tuned[key])
                nb = self._bytes(req)
                used += nb
                fresh.append(_Entry(req, m, nb, now))

        def collect():
            for e in to_collect:
                parcels = e.member.get_parcels() if e.status == 0 else None
                results[e.req.rid] = Result(e.req.rid, parcels, e.t, e.admitted, clock() - t0, e.req.arrival,
                                            e.steps, e.status, e.message)
                to_free.append(e)
            to_collect.clear()

        def release():
            nonlocal used
            for e in to_free:
                e.member.close()
                used -= e.nbytes
            to_free.clear()

        while waiting or cohorts or fresh or to_collect:
            if fresh:
                cohorts.append(_Cohort(fresh))
                fresh = []
            if not cohorts:
                collect()
                release()
                if not waiting:
                    break
                now = clock() - t0
                if waiting[0].arrival > now:
                    time.sleep(min(idle_sleep, waiting[0].arrival - now))
                admit(clock() - t0)
                continue
            # one tick: every cohort's step in flight together
            a0 = clock()
            for c in cohorts:
                c.batch.launch([e.t for e in c.entries], [e.req.dt for e in c.entries])
            # overlapping the tick: results of the last retirements, admissions
            collect()
            admit(clock() - t0)
            if fresh:
                cohorts_new = _Cohort(fresh)
                fresh = []
            else:
                cohorts_new = None
            stepped = 0
            nmem = 0
            retired_any = []
            ops = []
            for c in cohorts:
                ts, status = c.batch.wait()
                for e, t, st in zip(c.entries, ts, status):
                    nmem += 1
                    if st == 0:
                        e.t = float(t)
                        e.steps += 1
                        stepped += e.member.n
                        if (self.parcel_ops_every and e.steps % self.parcel_ops_every == 0
                                and e.steps < e.req.steps):
                            ops.append(e)
                    else:
                        e.status = int(st)
                        e.message = e.member.last_error()
            # the members' parcel operations run concurrently, one host
            # thread per member (each on its own stream; the library takes
            # one host thread per handle)
            o0 = clock()
            for e, err in zip(ops, pool.map(_parcel_ops, ops) if len(ops) > 1 else map(_parcel_ops, ops)):
                if err is not None:
                    e.status = int(err.code)
                    e.message = str(err)
            ops_ms.append((clock() - o0) * 1e3)
            for c in cohorts:
                keep = []
                for e in c.entries:
                    if e.status != 0 or e.steps >= e.req.steps:
                        to_collect.append(e)   # its memory counts until released
                    else:
                        keep.append(e)
                if len(keep) != len(c.entries):
                    retired_any.append((c, keep))
            tick_ms.append((clock() - a0) * 1e3)
            tick_parcels.append(stepped)
            tick_members.append(nmem)
            # no step in flight: release memory, rebuild the cohorts that lost members
            release()
            for c, keep in retired_any:
                cohorts.remove(c)
                c.close()
                if keep:
                    cohorts.append(_Cohort(keep))
            if cohorts_new is not None:
                cohorts.append(cohorts_new)
        wall = clock() - t0
        pool.shutdown()
        for c in cohorts:
            c.close()
        lat = np.array(tick_ms) if tick_ms else np.zeros(1)
        parcel_steps = float(np.sum(tick_parcels))
        busy = float(np.sum(tick_ms)) * 1e-3
        done = [r for r in results.values()]
        stats = {
            "ticks": len(tick_ms),
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            # the parcel operations' share of a tick (merge + split of the
            # members that stepped, on concurrent host threads)
            "ops_ms_p50": float(np.percentile(ops_ms, 50)) if ops_ms else 0.0,
            "ops_ms_mean": float(np.mean(ops_ms)) if ops_ms else 0.0,
            "parcel_steps": parcel_steps,
            "wall_s": wall,
            "parcel_steps_per_s_wall": parcel_steps / wall if wall > 0 else 0.0,
            "parcel_steps_per_s_busy": parcel_steps / busy if busy > 0 else 0.0,
            "mean_members_per_tick": float(np.mean(tick_members)) if tick_members else 0.0,
            "queue_wait_s_p50": float(np.percentile([r.admitted - r.arrival for r in done], 50)) if done else 0.0,
            "failed": sum(1 for r in done if r.status != 0),
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
