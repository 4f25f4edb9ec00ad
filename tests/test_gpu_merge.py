"""parcel_merge on the GPU against the FP64 oracle, through the C-ABI.

Both sides take the same FP32-rounded inputs.  Which parcels merge, into
which neighbour, and the order of the container after the delete are
integer decisions: the parcel count and order must match exactly.  Merged
attributes are FP64 sums on both sides, stored in FP32 on the GPU."""
import dataclasses

import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import Member
from tests._util import maxrel, oracle_state

pytestmark = pytest.mark.gpu


def _with_small(mem, frac, seed, s=0.03):
    rng = np.random.default_rng(seed)
    idx = rng.random(mem.n) < frac
    vol = mem.volume.copy()
    B = mem.B.copy()
    vol[idx] = vol[idx] * s
    B[idx] = B[idx] * s ** (2.0 / 3.0)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)   # noqa: E731
    return dataclasses.replace(mem, volume=f32(vol), B=f32(B)), int(idx.sum())


@pytest.mark.parametrize("frac", [0.02, 0.1, 0.3, 0.5])
def test_merge_matches_oracle(frac):
    mem, nsmall = _with_small(W.rising_thermal(16, seed=71, humidity_plume=True, n_blobs=2), frac, 72)
    print(f"merge: thermal16, {nsmall} small parcels", flush=True)
    gm = Member.from_workload(mem)
    nm = gm.parcel_merge()
    st = oracle_state(mem)
    nm_ref = st.parcel_merge()
    assert nm == nm_ref and gm.n == st.n
    # collect_merge_stats: n_big_close (the diagnostics' nbig_iclo) and the
    # n-way merger counts
    assert gm.parcel_diagnostics()["nbig_iclo"] == st.merge_stats[0]
    assert gm.parcel_merge_stats().tolist() == st.merge_stats[1:].tolist()
    p = gm.get_parcels()
    assert np.abs(p["position"] - st.position).max() <= 1e-12 * max(mem.extent)
    assert maxrel(p["volume"], st.volume) <= 1e-6
    assert maxrel(p["B"], st.B) <= 1e-6
    assert maxrel(p["vorticity"], st.vorticity) <= 1e-6
    assert maxrel(p["buoyancy"], st.buoyancy) <= 1e-6
    assert maxrel(p["humidity"], st.humidity) <= 1e-6


def test_merge_then_split_then_step():
    mem, _ = _with_small(W.rising_thermal(16, seed=73), 0.05, 74)
    gm = Member.from_workload(mem, max_parcels=2 * mem.n)
    gm.parcel_merge()
    gm.parcel_split(lambda_max=3.0)
    gm.ls_rk_step(0.0, 0.01)
    st = oracle_state(mem)
    st.parcel_merge()
    st.parcel_split(lambda_max=3.0)
    st.ls_rk_step(0.0, 0.01)
    p = gm.get_parcels()
    assert gm.n == st.n
    assert np.abs(p["position"] - st.position).max() / max(mem.dx) <= 1e-5
    assert maxrel(p["vorticity"], st.vorticity) <= 1e-5


def test_step_graph_follows_parcel_count_changes():
    """A member keeps its captured step graph across merges and splits (the
    parcel count changes, the launch sequence does not; only the argument
    tables are replaced): each later step gives the same bits as a fresh
    member holding the same parcels, and so does a batch of two members
    whose counts both change between its steps."""
    mem, _ = _with_small(W.rising_thermal(16, seed=75), 0.05, 76)
    a = Member.from_workload(mem, max_parcels=2 * mem.n)
    t = a.ls_rk_step(0.0, 0.01)
    for step in range(3):
        nm = a.parcel_merge()
        ns = a.parcel_split(lambda_max=3.0)
        p = a.get_parcels()
        b = Member.from_workload(dataclasses.replace(mem, **p), max_parcels=2 * mem.n)
        print(f"step {step}: merged {nm}, split {ns}, n {a.n}", flush=True)
        ta = a.ls_rk_step(t, 0.01, t_initial=t)
        tb = b.ls_rk_step(t, 0.01, t_initial=t)
        assert ta == tb
        pa, pb = a.get_parcels(), b.get_parcels()
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), (step, k)
        b.close()
        t = ta
    assert a.n != mem.n

    from epic_amd.core import Batch

    w1 = _with_small(W.rising_thermal(16, ny=24, seed=77), 0.05, 78)[0]
    ws = (mem, w1)
    ms = [Member.from_workload(w, max_parcels=2 * w.n) for w in ws]
    bt = Batch(ms)
    tt = np.zeros(2)
    for step in range(3):
        tt, st = bt.step(tt, [0.01, 0.01], t_initial=tt)
        assert (st == 0).all()
        for m in ms:
            m.parcel_merge()
            m.parcel_split(lambda_max=3.0)
        refs = [Member.from_workload(dataclasses.replace(w, **m.get_parcels()), max_parcels=2 * w.n)
                for w, m in zip(ws, ms)]
        t0 = tt.copy()
        tt, st = bt.step(tt, [0.01, 0.01], t_initial=tt)
        assert (st == 0).all()
        for i, (m, r) in enumerate(zip(ms, refs)):
            assert r.ls_rk_step(t0[i], 0.01, t_initial=t0[i]) == tt[i]
            got, want = m.get_parcels(), r.get_parcels()
            for k in got:
                assert np.array_equal(got[k], want[k]), (step, i, k)
            r.close()
    bt.close()
