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
