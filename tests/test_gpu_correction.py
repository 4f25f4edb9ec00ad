"""Parcel correction on the GPU against the FP64 oracle, through the C-ABI
(parcel_correction.f90: apply_laplace, apply_gradient, init_parcel_correction,
apply_vortcor).

Same FP32-rounded inputs on both sides.  The displacements come from
volg (FP32 on the GPU) through FP32 horizontal FFTs and the FP64 z solve;
the bar is max|d_gpu - d_ref| <= 1e-5 max|d_ref| per correction, the field
tolerance of the per-timestep core (positions themselves stay FP64).
apply_vortcor shifts FP32 vorticity: the bar is 1e-6 relative to max|vor|.
"""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import Member
from tests._util import maxrel, oracle_state

pytestmark = pytest.mark.gpu

TOL_DISP = 1e-5


def _displacement(after, before, extent):
    d = after - before
    for k in range(2):   # periodic in x and y
        d[:, k] -= extent[k] * np.round(d[:, k] / extent[k])
    return d


def _thermal(n, seed, ny=None, nz=None):
    mem = W.rising_thermal(n, seed=seed, ny=ny, nz=nz, humidity_plume=True, n_blobs=2)
    # uneven volumes so that phi = volg / vcell - 1 is far from zero
    rng = np.random.default_rng(seed)
    s = (1.0 + 0.3 * rng.uniform(-1.0, 1.0, mem.n)).astype(np.float32).astype(np.float64)
    mem.volume = (mem.volume * s).astype(np.float32).astype(np.float64)
    mem.B = (mem.B * s[:, None] ** (2.0 / 3.0)).astype(np.float32).astype(np.float64)
    return mem


@pytest.mark.parametrize("shape", [(16, None, None), (32, 24, 16), (24, 32, 12)])
def test_laplace_matches_oracle(shape):
    n, ny, nz = shape
    mem = _thermal(n, 91, ny, nz)
    print(f"apply_laplace: thermal {mem.nx}x{mem.ny}x{mem.nz} n={mem.n}", flush=True)
    gm = Member.from_workload(mem)
    gm.apply_laplace()
    st = oracle_state(mem)
    st.apply_laplace()
    d_ref = _displacement(st.position, mem.position, mem.extent)
    d_gpu = _displacement(gm.get_parcels()["position"], mem.position, mem.extent)
    err = np.abs(d_gpu - d_ref).max() / np.abs(d_ref).max()
    print(f"  max|d_ref| = {np.abs(d_ref).max():.3e} dx, rel err {err:.2e}", flush=True)
    assert np.abs(d_ref).max() > 1e-3 * max(mem.dx)
    assert err <= TOL_DISP
    # x and y stay in the periodic box
    pos = gm.get_parcels()["position"]
    for k in range(2):
        assert pos[:, k].min() >= mem.lower[k] - 1e-12
        assert pos[:, k].max() <= mem.lower[k] + mem.extent[k] + 1e-12


@pytest.mark.parametrize("shape", [(16, None, None), (32, 24, 16)])
def test_gradient_matches_oracle(shape):
    n, ny, nz = shape
    mem = _thermal(n, 92, ny, nz)
    print(f"apply_gradient: thermal {mem.nx}x{mem.ny}x{mem.nz} n={mem.n}", flush=True)
    gm = Member.from_workload(mem)
    gm.apply_gradient(1.8, 0.5)
    st = oracle_state(mem)
    st.apply_gradient(1.8, 0.5)
    d_ref = _displacement(st.position, mem.position, mem.extent)
    d_gpu = _displacement(gm.get_parcels()["position"], mem.position, mem.extent)
    err = np.abs(d_gpu - d_ref).max() / np.abs(d_ref).max()
    print(f"  max|d_ref| = {np.abs(d_ref).max():.3e}, rel err {err:.2e}", flush=True)
    assert err <= TOL_DISP


def test_correction_iterations_of_the_run_loop():
    """epic3d.f90:148-151: correction_iters = 2 rounds of apply_laplace
    (reusing the support points after the first) and apply_gradient."""
    mem = _thermal(16, 93)
    print(f"correction loop: thermal16 n={mem.n}", flush=True)
    gm = Member.from_workload(mem)
    st = oracle_state(mem)
    for it in (1, 2):
        gm.apply_laplace(l_reuse=it > 1)
        gm.apply_gradient(1.8, 0.5, l_reuse=True)
        st.apply_laplace(l_reuse=it > 1)
        st.apply_gradient(1.8, 0.5, l_reuse=True)
    d_ref = _displacement(st.position, mem.position, mem.extent)
    d_gpu = _displacement(gm.get_parcels()["position"], mem.position, mem.extent)
    err = np.abs(d_gpu - d_ref).max() / np.abs(d_ref).max()
    print(f"  rel err {err:.2e}", flush=True)
    assert err <= 4 * TOL_DISP
    # the correction evens out the gridded volume
    gm.par2grid()
    st.par2grid()
    assert maxrel(gm.get_field("volg"), st.volg) <= 1e-5


def test_step_after_correction_matches_oracle():
    mem = _thermal(16, 94)
    gm = Member.from_workload(mem)
    st = oracle_state(mem)
    gm.apply_laplace()
    st.apply_laplace()
    gm.ls_rk_step(0.0, 0.01)
    st.ls_rk_step(0.0, 0.01)
    p = gm.get_parcels()
    assert np.abs(p["position"] - st.position).max() / max(mem.dx) <= 1e-5
    assert maxrel(p["vorticity"], st.vorticity) <= 1e-5


def test_vortcor_matches_oracle():
    mem = _thermal(16, 95)
    gm = Member.from_workload(mem)
    st = oracle_state(mem)
    vb = gm.init_parcel_correction()
    vb_ref = st.volume_weighted_vorticity()
    assert np.allclose(vb, vb_ref, rtol=1e-12, atol=1e-12 * np.abs(mem.vorticity).max())
    # a step changes the net vorticity (buoyancy torque); vortcor resets it
    gm.ls_rk_step(0.0, 0.02)
    st.ls_rk_step(0.0, 0.02)
    gm.apply_vortcor()
    st.apply_vortcor(vb_ref)
    p = gm.get_parcels()
    assert maxrel(p["vorticity"], st.vorticity) <= 1e-5
    w = (p["volume"][:, None] * p["vorticity"]).sum(0) / p["volume"].sum()
    assert np.abs(w - vb_ref).max() <= 1e-6 * np.abs(p["vorticity"]).max()


def test_empty_member_is_a_no_op():
    mem = W.rising_thermal(16, seed=96)
    gm = Member.from_workload(mem)
    gm.set_parcels(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 5)), np.zeros(0), np.zeros(0), np.zeros(0))
    gm.apply_laplace()
    gm.apply_gradient()
    assert np.array_equal(gm.init_parcel_correction(), np.zeros(3))
    gm.apply_vortcor()
