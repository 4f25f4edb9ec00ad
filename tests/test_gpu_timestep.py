"""The adaptive time step on the GPU against the FP64 oracle, through the
C-ABI (get_time_step, rk_utils.f90:73-183, and ls_rk_step with it,
ls_rk.f90:88-136).

Same FP32-rounded inputs on both sides.  gmax and bmax are maxima over FP32
node fields that agree with the oracle's within 1e-5 max-rel, so dt agrees
within 1e-5 relative; the adaptive step then compares like the fixed step
(positions within 1e-5 of dx, attributes within 1e-5)."""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import EpicError, Member
from tests._util import maxrel, oracle_state

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("name", ["thermal16", "moist16", "beltrami16"])
def test_get_time_step_matches_oracle(name):
    mem = {"thermal16": lambda: W.rising_thermal(16, seed=81),
           "moist16": lambda: W.moist_member(16, seed=82),
           "beltrami16": lambda: W.beltrami(16)}[name]()
    print(f"get_time_step: {name}", flush=True)
    gm = Member.from_workload(mem)
    gm.par2grid()
    gm.vor2vel()
    st = oracle_state(mem)
    st.par2grid()
    st.vor2vel()
    dt, g, b = gm.get_time_step(0.0, alpha=0.2)
    dt_r, g_r, b_r = st.get_time_step(0.0, alpha=0.2)
    print(f"  dt {dt:.6e} / {dt_r:.6e}, gmax {g:.6e} / {g_r:.6e}, bmax {b:.6e} / {b_r:.6e}", flush=True)
    assert abs(g - g_r) <= TOL * g_r
    assert abs(b - b_r) <= TOL * b_r
    assert abs(dt - dt_r) <= TOL * dt_r
    # precise_stop and the fixed branch
    dt2, _, _ = gm.get_time_step(1.0 - 0.5 * dt_r, alpha=0.2, limit=1.0, precise_stop=True)
    assert dt2 == pytest.approx(0.5 * dt_r, rel=1e-9)
    assert gm.get_time_step(0.0, fixed_dt=0.004)[0] == 0.004
    with pytest.raises(EpicError) as e:
        gm.get_time_step(2.0, alpha=0.2, limit=1.0, precise_stop=True)
    assert e.value.code == -7


def test_adaptive_steps_match_oracle():
    mem = W.rising_thermal(16, seed=83, humidity_plume=True, n_blobs=2)
    print("adaptive steps: thermal16, three steps", flush=True)
    gm = Member.from_workload(mem)
    st = oracle_state(mem)
    t_g = t_r = 0.0
    for step in range(3):
        t_g, dt_g = gm.ls_rk_step_adaptive(t_g, alpha=0.2)
        t_r, dt_r = st.ls_rk_step_adaptive(t_r, alpha=0.2)
        print(f"  step {step}: dt {dt_g:.6e} / {dt_r:.6e}", flush=True)
        assert abs(dt_g - dt_r) <= TOL * dt_r
    p = gm.get_parcels()
    assert np.abs(p["position"] - st.position).max() / max(mem.dx) <= TOL
    assert maxrel(p["vorticity"], st.vorticity) <= TOL
    assert maxrel(p["B"], st.B) <= TOL


def test_adaptive_step_stops_at_the_limit_and_fixed_steps_still_work():
    mem = W.rising_thermal(16, seed=84)
    gm = Member.from_workload(mem)
    t, dt = gm.ls_rk_step_adaptive(0.0, alpha=0.2)
    t2, dt2 = gm.ls_rk_step_adaptive(t, alpha=0.2, limit=t + 0.25 * dt, precise_stop=True)
    assert dt2 == pytest.approx(0.25 * dt, rel=1e-9) and t2 == pytest.approx(t + 0.25 * dt, rel=1e-12)
    # the fixed-dt graph is a separate plan: interleaving keeps both correct
    t3 = gm.ls_rk_step(t2, 0.01)
    assert t3 == pytest.approx(t2 + 0.01, rel=1e-14)
    st = oracle_state(mem)
    tr, _ = st.ls_rk_step_adaptive(0.0, alpha=0.2)
    tr, _ = st.ls_rk_step_adaptive(tr, alpha=0.2, limit=tr + 0.25 * dt, precise_stop=True)
    st.ls_rk_step(tr, 0.01, t_initial=0.0)
    p = gm.get_parcels()
    assert np.abs(p["position"] - st.position).max() / max(mem.dx) <= TOL
    assert maxrel(p["vorticity"], st.vorticity) <= TOL
