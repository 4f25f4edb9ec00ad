"""Served ensemble: continuous batching under an HBM budget, through the C-ABI.

Requests of different grids, parcel counts, step counts and time steps
arrive over time; the budget admits only some at once, so members join and
leave a running batch.  Every result must equal the same member run alone
with the plain ls_rk_step, bit for bit (the batch runs each member's own
deterministic kernels)."""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import Member, member_bytes
from epic_amd.serve import Request, Server, poisson_requests

pytestmark = pytest.mark.gpu


def _alone(req):
    m = Member.from_workload(req.workload, rk_order=req.rk_order)
    t = 0.0
    for _ in range(req.steps):
        t = m.ls_rk_step(t, req.dt)
    p = m.get_parcels()
    m.close()
    return p, t


def test_served_results_equal_alone():
    ws = [W.rising_thermal(16, seed=31), W.rising_thermal(32, ny=16, nz=24, seed=32), W.beltrami(16),
          W.rising_thermal(16, seed=33, humidity_plume=True, n_blobs=2), W.rising_thermal(24, ny=16, nz=16, seed=34)]
    reqs = [Request(0, ws[0], 3, 0.01, 0.0), Request(1, ws[1], 2, 0.01, 0.0), Request(2, ws[2], 4, 0.02, 0.05),
            Request(3, ws[3], 2, 0.005, 0.05, rk_order=3), Request(4, ws[4], 3, 0.01, 0.1)]
    sizes = [member_bytes(w.nx, w.ny, w.nz, w.n) for w in ws]
    # room for the two largest members at once: later requests wait for memory
    budget = sum(sorted(sizes)[-2:])
    print(f"serve: {len(reqs)} requests, budget {budget / 2**20:.1f} MiB", flush=True)
    results, stats = Server(budget, max_members=3).run(reqs)
    print(f"serve: stats {stats}", flush=True)
    assert sorted(results) == [0, 1, 2, 3, 4]
    assert stats["ticks"] >= max(r.steps for r in reqs)
    for r in reqs:
        p, t = _alone(r)
        got = results[r.rid]
        assert abs(got.t - t) < 1e-15
        for k in ("position", "vorticity", "B", "volume"):
            assert np.array_equal(got.parcels[k], p[k]), (r.rid, k)


def _alone_with_parcel_ops(req, every, capacity):
    m = Member.from_workload(req.workload, rk_order=req.rk_order, max_parcels=capacity)
    t = 0.0
    for s in range(1, req.steps + 1):
        t = m.ls_rk_step(t, req.dt)
        if s % every == 0 and s < req.steps:   # epic3d.f90:143-148
            m.parcel_merge()
            m.parcel_split()
    p = m.get_parcels()
    m.close()
    return p, t


def test_served_parcel_operations_equal_alone():
    """Merge and split every second step of each member (the run loop's
    parcel operations): members change their parcel counts, their cohort is
    captured again, and every result still equals the member run alone."""
    ws = [W.rising_thermal(16, seed=51), W.rising_thermal(24, ny=16, nz=16, seed=52), W.rising_thermal(16, ny=24, seed=53)]
    reqs = [Request(0, ws[0], 5, 0.01, 0.0), Request(1, ws[1], 4, 0.01, 0.0), Request(2, ws[2], 6, 0.01, 0.02)]
    results, stats = Server(1 << 34, max_members=2, parcel_ops_every=2).run(reqs)
    assert sorted(results) == [0, 1, 2] and stats["failed"] == 0
    grown = 0
    for r in reqs:
        p, t = _alone_with_parcel_ops(r, 2, 2 * r.workload.n)
        got = results[r.rid]
        assert got.status == 0, got.message
        assert abs(got.t - t) < 1e-15
        grown += got.parcels["volume"].shape[0] != r.workload.n
        for k in ("position", "vorticity", "B", "volume"):
            assert np.array_equal(got.parcels[k], p[k]), (r.rid, k)
    assert grown > 0   # the splits changed parcel counts


def test_poisson_arrivals_and_autotuned_tiles():
    ws = [W.rising_thermal(16, seed=40 + i) for i in range(4)]
    reqs = poisson_requests(ws, rate=50.0, steps=(1, 3), dt=0.01, seed=7)
    results, stats = Server(1 << 34, max_members=8, autotune=True).run(reqs)
    assert sorted(results) == [r.rid for r in reqs]
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    assert stats["parcel_steps"] == sum(r.workload.n * r.steps for r in reqs)


def test_failed_member_is_retired_with_its_status():
    """A member whose step fails (a parcel more than one cell below the
    lower wall: "Wrong total number of parcels", parcel_interpl.f90:248-250)
    leaves the batch with its status and message; the others finish and
    equal their runs alone."""
    import dataclasses

    good = [W.rising_thermal(16, seed=51), W.rising_thermal(16, ny=24, nz=12, seed=52)]
    bad = W.rising_thermal(16, seed=53)
    pos = bad.position.copy()
    pos[17, 2] = bad.lower[2] - 1.5 * bad.dx[2]
    bad = dataclasses.replace(bad, position=pos)
    reqs = [Request(0, good[0], 3, 0.01, 0.0), Request(1, bad, 3, 0.01, 0.0), Request(2, good[1], 2, 0.01, 0.0)]
    results, stats = Server(1 << 34, max_members=8).run(reqs)
    assert results[1].status == -3 and results[1].parcels is None and results[1].steps_done == 0
    assert "Wrong total number of parcels" in results[1].message
    assert stats["failed"] == 1
    for r in (reqs[0], reqs[2]):
        p, t = _alone(r)
        got = results[r.rid]
        assert got.status == 0 and got.steps_done == r.steps and abs(got.t - t) < 1e-15
        for k in ("position", "vorticity", "B", "volume"):
            assert np.array_equal(got.parcels[k], p[k]), (r.rid, k)


def test_launch_wait_overlaps_admission():
    """Two batches in flight together, a third member created and uploaded
    while they run: each result equals its run alone."""
    from epic_amd.core import Batch

    ws = [W.rising_thermal(16, seed=61), W.rising_thermal(24, ny=16, nz=16, seed=62), W.beltrami(16)]
    ms = [Member.from_workload(w) for w in ws[:2]]
    b1, b2 = Batch([ms[0]]), Batch([ms[1]])
    t1 = t2 = np.zeros(1)
    late = None
    for step in range(3):
        b1.launch(t1, 0.01)
        b2.launch(t2, 0.01)
        if step == 0:
            late = Member.from_workload(ws[2])
        t1, s1 = b1.wait()
        t2, s2 = b2.wait()
        assert s1[0] == 0 and s2[0] == 0
    tl = 0.0
    for _ in range(3):
        tl = late.ls_rk_step(tl, 0.01)
    for w, m in zip(ws, ms + [late]):
        p, _ = _alone(Request(0, w, 3, 0.01, 0.0))
        got = m.get_parcels()
        for k in ("position", "vorticity", "B"):
            assert np.array_equal(got[k], p[k]), (w.name, k)
    b1.close()
    b2.close()
