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


def test_poisson_arrivals_and_autotuned_tiles():
    ws = [W.rising_thermal(16, seed=40 + i) for i in range(4)]
    reqs = poisson_requests(ws, rate=50.0, steps=(1, 3), dt=0.01, seed=7)
    results, stats = Server(1 << 34, max_members=8, autotune=True).run(reqs)
    assert sorted(results) == [r.rid for r in reqs]
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    assert stats["parcel_steps"] == sum(r.workload.n * r.steps for r in reqs)
