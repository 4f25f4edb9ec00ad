"""Physical re-sort of the parcel arrays into cell order is not observable:
par2grid sums are exact integers per tile and every other kernel works per
parcel, so steps with a re-sort after every step give bitwise the state of
steps without one, and epic_get_parcels returns the caller's order."""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import Member

pytestmark = pytest.mark.gpu


def _state(m):
    p = m.get_parcels()
    return np.concatenate([p[k].ravel() for k in sorted(p)])


def test_resort_is_not_observable():
    mem = W.rising_thermal(32, ny=16, nz=24, seed=95)
    # scramble the caller order so that the physical order is far from the cells
    perm = np.random.default_rng(0).permutation(mem.n)
    for k in ("position", "vorticity", "B", "volume", "buoyancy", "humidity"):
        setattr(mem, k, getattr(mem, k)[perm])
    print(f"resort: thermal32x16x24 n={mem.n}", flush=True)
    a = Member.from_workload(mem)
    b = Member.from_workload(mem)
    a.set_resort_interval(1)
    b.set_resort_interval(0)
    a.resort()
    assert np.array_equal(_state(a), _state(b))
    t_a = t_b = 0.0
    for _ in range(3):
        t_a = a.ls_rk_step(t_a, 0.01)
        t_b = b.ls_rk_step(t_b, 0.01)
    assert np.array_equal(_state(a), _state(b))
    for f in ("volg", "vortg", "velog"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f


def _with_small(mem, frac, seed, s=0.03):
    import dataclasses

    rng = np.random.default_rng(seed)
    idx = rng.random(mem.n) < frac
    vol = mem.volume.copy()
    B = mem.B.copy()
    vol[idx] = vol[idx] * s
    B[idx] = B[idx] * s ** (2.0 / 3.0)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)   # noqa: E731
    return dataclasses.replace(mem, volume=f32(vol), B=f32(B))


def test_merge_and_split_after_resorts_follow_caller_order():
    """Merge pairings, the hole filling and the child order follow the
    caller's index, not the physical (cell) order a re-sort leaves: after
    steps with a re-sort every step, merge and split give bitwise the state
    of the same steps without re-sorts, and the merge equals the oracle's on
    the same parcels."""
    from tests._util import maxrel, oracle_state

    mem = _with_small(W.rising_thermal(16, ny=24, nz=12, seed=96), 0.08, 97)
    a = Member.from_workload(mem, max_parcels=2 * mem.n)
    b = Member.from_workload(mem, max_parcels=2 * mem.n)
    a.set_resort_interval(1)
    b.set_resort_interval(0)
    t_a = t_b = 0.0
    for _ in range(3):
        t_a = a.ls_rk_step(t_a, 0.01)
        t_b = b.ls_rk_step(t_b, 0.01)
    assert np.array_equal(_state(a), _state(b))
    before = a.get_parcels()
    nm_a, nm_b = a.parcel_merge(), b.parcel_merge()
    assert nm_a == nm_b > 0
    assert np.array_equal(_state(a), _state(b))
    # the oracle merges the same parcels (the device's stored values)
    import dataclasses

    st = oracle_state(dataclasses.replace(mem, **before))
    assert st.parcel_merge() == nm_a
    p = a.get_parcels()
    assert a.n == st.n
    assert np.abs(p["position"] - st.position).max() <= 1e-12 * max(mem.extent)
    for k in ("volume", "B", "vorticity", "buoyancy"):
        assert maxrel(p[k], getattr(st, k)) <= 1e-6, k
    # split after a further re-sorted step
    t_a = a.ls_rk_step(t_a, 0.01)
    t_b = b.ls_rk_step(t_b, 0.01)
    ns_a, ns_b = a.parcel_split(lambda_max=2.5), b.parcel_split(lambda_max=2.5)
    assert ns_a == ns_b > 0
    assert np.array_equal(_state(a), _state(b))
    for x, y in zip(a.get_tendencies(), b.get_tendencies()):
        assert np.array_equal(x, y)
    # merge and split leave the slots out of caller order (relabelled caller
    # indices, children at the end): further re-sorted steps and parcel
    # operations still give the same caller-ordered state
    for cycle in range(2):
        for _ in range(2):
            t_a = a.ls_rk_step(t_a, 0.01)
            t_b = b.ls_rk_step(t_b, 0.01)
        assert np.array_equal(_state(a), _state(b)), cycle
        assert a.parcel_merge() == b.parcel_merge()
        for x, y in zip(a.get_tendencies(), b.get_tendencies()):
            assert np.array_equal(x, y), cycle
        assert a.parcel_split(lambda_max=2.5) == b.parcel_split(lambda_max=2.5)
        assert np.array_equal(_state(a), _state(b)), cycle


def test_tendencies_follow_their_parcels_across_resort():
    mem = W.rising_thermal(16, seed=98)
    perm = np.random.default_rng(3).permutation(mem.n)
    for k in ("position", "vorticity", "B", "volume", "buoyancy", "humidity"):
        setattr(mem, k, getattr(mem, k)[perm])
    a = Member.from_workload(mem)
    b = Member.from_workload(mem)
    for m in (a, b):
        m.par2grid()
        m.vor2vel()
        m.vorticity_tendency()
        m.grid2par()
    ta = a.get_tendencies()
    a.resort()
    for x, y in zip(ta, a.get_tendencies()):
        assert np.array_equal(x, y)
    # grid2par(add) onto re-sorted tendencies adds onto each parcel's own
    a.resort()
    a.grid2par(add=True)
    b.grid2par(add=True)
    for x, y in zip(a.get_tendencies(), b.get_tendencies()):
        assert np.array_equal(x, y)
    a.resort()
    da, db = a.parcel_diagnostics(), b.parcel_diagnostics()
    for k in ("ke", "enstrophy", "sum_vol", "avg_lam"):
        assert abs(da[k] - db[k]) <= 1e-12 * max(1.0, abs(db[k])), k
