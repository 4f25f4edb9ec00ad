"""Batched LS-RK step and tiling autotuner, through the C-ABI.

A batch runs every kernel of the step as one launch over all its members,
captured as one HIP graph.  Each member's kernels and arguments are the same
as in its plain step, and every sum is formed in a fixed order, so a member
advanced in a batch must match the same member advanced alone bit for bit,
whatever the other members are.  The autotuner may change the tile z chunk,
which regroups the par2grid partial sums: after tuning the step stays within
the FP64-oracle tolerance, and tuning itself leaves the parcel state unchanged.
"""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import Batch, Member
from tests._util import maxrel, oracle_state

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _mix():
    return [
        (W.rising_thermal(16, seed=21), 4, 0.01),
        (W.rising_thermal(32, ny=16, nz=24, seed=22), 4, 0.01),
        (W.beltrami(16), 3, 0.02),
        (W.rising_thermal(16, seed=23, humidity_plume=True, n_blobs=4), 4, 0.005),
    ]


def _state(m):
    p = m.get_parcels()
    return np.concatenate([p["position"].ravel(), p["vorticity"].ravel(), p["B"].ravel()])


def test_batch_equals_plain_bitwise():
    mix = _mix()
    print(f"batch: {len(mix)} members {[w.name for w, _, _ in mix]}", flush=True)
    alone = [Member.from_workload(w, rk_order=o) for w, o, _ in mix]
    inmix = [Member.from_workload(w, rk_order=o) for w, o, _ in mix]
    b = Batch(inmix)
    t_alone = [0.0] * len(mix)
    t_mix = np.zeros(len(mix))
    dts = np.array([d for _, _, d in mix])
    for step in range(3):
        for i, (m, (_, _, d)) in enumerate(zip(alone, mix)):
            t_alone[i] = m.ls_rk_step(t_alone[i], d)
        t_mix, st = b.step(t_mix, dts)
        assert (st == 0).all()
    assert np.allclose(t_mix, t_alone, rtol=0, atol=1e-15)
    for i in range(len(mix)):
        assert np.array_equal(_state(alone[i]), _state(inmix[i])), mix[i][0].name
        ga, gb = alone[i], inmix[i]
        ga.par2grid()
        gb.par2grid()
        assert np.array_equal(ga.get_field("volg"), gb.get_field("volg"))
    b.close()


def test_batch_subset_and_order_do_not_matter():
    """A member's result does not depend on which other members share the batch."""
    mix = _mix()
    a = [Member.from_workload(w, rk_order=o) for w, o, _ in mix]
    b = [Member.from_workload(w, rk_order=o) for w, o, _ in mix]
    dts = np.array([d for _, _, d in mix])
    ba = Batch(a)
    bb = Batch([b[2], b[0]])
    bc = Batch([b[3], b[1]])
    ta, _ = ba.step(np.zeros(4), dts)
    tb, _ = bb.step(np.zeros(2), dts[[2, 0]])
    tc, _ = bc.step(np.zeros(2), dts[[3, 1]])
    for i in range(4):
        assert np.array_equal(_state(a[i]), _state(b[i])), i
    for x in (ba, bb, bc):
        x.close()


def test_batch_step_matches_oracle():
    w, order, dt = _mix()[1]
    m = Member.from_workload(w, rk_order=order)
    b = Batch([m])
    b.step([0.0], [dt])
    st = oracle_state(w)
    st.ls_rk_step(0.0, dt)
    p = m.get_parcels()
    assert np.abs(p["position"] - st.position).max() / max(w.dx) <= TOL
    assert maxrel(p["vorticity"], st.vorticity) <= TOL
    assert maxrel(p["B"], st.B) <= TOL
    b.close()


def test_autotune_leaves_state_and_stays_accurate():
    w = W.rising_thermal(32, ny=16, nz=24, seed=24)
    m = Member.from_workload(w)
    before = _state(m)
    tz = m.autotune()
    print(f"autotune: tz_p2g={tz[0]} tz_g2p={tz[1]}", flush=True)
    assert tz[0] in (2, 4, 8) and tz[1] in (2, 4, 8)
    assert np.array_equal(before, _state(m))
    m.ls_rk_step(0.0, 0.01)
    st = oracle_state(w)
    st.ls_rk_step(0.0, 0.01)
    p = m.get_parcels()
    assert np.abs(p["position"] - st.position).max() / max(w.dx) <= TOL
    assert maxrel(p["vorticity"], st.vorticity) <= TOL
    assert maxrel(p["B"], st.B) <= TOL


@pytest.mark.parametrize("tz", [(4, 2), (8, 4), (2, 8)])
def test_every_tiling_matches_oracle(tz):
    w = W.rising_thermal(16, seed=25)
    m = Member.from_workload(w)
    m.set_tiling(*tz)
    m.par2grid()
    st = oracle_state(w)
    st.par2grid()
    for f in ("volg", "vortg", "tbuoyg"):
        got, ref = m.get_field(f), getattr(st, f)
        assert maxrel(got[..., 1:17, 1:17, :], ref[..., 1:17, 1:17, :]) <= TOL, (tz, f)
    m2 = Member.from_workload(w)
    m2.set_tiling(*tz)
    m2.ls_rk_step(0.0, 0.01)
    st2 = oracle_state(w)
    st2.ls_rk_step(0.0, 0.01)
    p = m2.get_parcels()
    assert np.abs(p["position"] - st2.position).max() / max(w.dx) <= TOL
    assert maxrel(p["vorticity"], st2.vorticity) <= TOL
