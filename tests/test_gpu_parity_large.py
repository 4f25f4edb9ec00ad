"""GPU parity against the FP64 oracle at the workloads' own sizes, through
the C-ABI.

- The rising thermal at 128^3 (16,777,216 parcels): one par2grid, vor2vel,
  vorticity_tendency and grid2par from the same inputs, then one full LS-RK4
  step; every gridded field and parcel attribute within 1e-5 max-relative
  error of FP64 (positions relative to the cell size), parcel counts
  bit-exact.
- 256^3 fields only: a k^-3 Gaussian random vorticity field and 32 buoyancy
  blobs set on the grid, vor2vel and vorticity_tendency within 1e-5
  (the FP32 horizontal FFTs at their longest lengths).

Each field's error is printed, so the log of a run is the error table.
"""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import Member
from tests._util import maxrel, oracle_state

pytestmark = pytest.mark.gpu
TOL = 1e-5


def interior(a, nx, ny):
    return a[..., 1:nx + 1, 1:ny + 1, :]


_cache = {}


def _thermal128():
    if "t128" not in _cache:
        print("  generating the 128^3 rising thermal", flush=True)
        _cache["t128"] = W.rising_thermal(128)
    return _cache["t128"]


def _report(tag, name, e):
    print(f"  {tag} {name}: max-rel {e:.2e}", flush=True)


def test_thermal128_operators():
    mem = _thermal128()
    nx, ny = mem.nx, mem.ny
    st = oracle_state(mem)
    gm = Member.from_workload(mem)
    st.par2grid()
    gm.par2grid()
    for f in ("volg", "vortg", "tbuoyg", "dbuoyg", "humg"):
        e = maxrel(interior(gm.get_field(f), nx, ny), interior(getattr(st, f), nx, ny))
        _report("128^3 par2grid", f, e)
        assert e <= TOL, (f, e)
    for f in ("nparg", "nsparg"):
        assert np.array_equal(interior(gm.get_field(f), nx, ny), interior(getattr(st, f), nx, ny)), f
    # the inversions from the oracle's fields (same input on both sides)
    gi = Member(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, mem.glat, mem.lambda_c, mem.q0, mem.f_cor,
                mem.zeta_zero, max_parcels=1)
    gi.set_field("vortg", st.vortg)
    gi.set_field("tbuoyg", st.tbuoyg)
    st.vor2vel()
    gi.vor2vel()
    for f in ("vortg", "velog", "velgradg"):
        e = maxrel(interior(gi.get_field(f), nx, ny), interior(getattr(st, f), nx, ny))
        _report("128^3 vor2vel", f, e)
        assert e <= TOL, (f, e)
    st.vorticity_tendency()
    gi.vorticity_tendency()
    e = maxrel(interior(gi.get_field("vtend"), nx, ny), interior(st.vtend, nx, ny))
    _report("128^3 vorticity_tendency", "vtend", e)
    assert e <= TOL, e
    gi.close()
    # grid2par from the oracle's fields
    for f in ("velog", "velgradg", "vtend"):
        gm.set_field(f, getattr(st, f))
    st.grid2par()
    gm.grid2par()
    dp, dv, s = gm.get_tendencies()
    for got, ref, f in ((dp, st.delta_pos, "delta_pos"), (dv, st.delta_vor, "delta_vor"), (s, st.strain, "strain")):
        e = maxrel(got, ref)
        _report("128^3 grid2par", f, e)
        assert e <= TOL, (f, e)
    gm.close()


def test_thermal128_rk4_step():
    mem = _thermal128()
    nx, ny = mem.nx, mem.ny
    st = oracle_state(mem)
    gm = Member.from_workload(mem)
    t_ref = st.ls_rk_step(0.0, 0.01)
    t_gpu = gm.ls_rk_step(0.0, 0.01)
    assert t_ref == t_gpu
    got = gm.get_parcels()
    e = np.abs(got["position"] - st.position).max() / max(mem.dx)
    _report("128^3 RK4 step", "position (/dx)", e)
    assert e <= TOL, e
    for f in ("vorticity", "B", "volume", "buoyancy", "humidity"):
        e = maxrel(got[f], getattr(st, f))
        _report("128^3 RK4 step", f, e)
        assert e <= TOL, (f, e)
    for f in ("volg", "vortg", "tbuoyg", "velog", "velgradg", "vtend"):
        e = maxrel(interior(gm.get_field(f), nx, ny), interior(getattr(st, f), nx, ny))
        _report("128^3 RK4 step (last stage)", f, e)
        assert e <= TOL, (f, e)
    for f in ("nparg", "nsparg"):
        assert np.array_equal(interior(gm.get_field(f), nx, ny), interior(getattr(st, f), nx, ny)), f
    gm.close()


def test_fields256_vor2vel_vtend():
    n = 256
    rng = np.random.default_rng(20261202)
    ext = np.pi
    lower, extent = (-0.5 * ext,) * 3, (ext,) * 3
    print("  generating 256^3 vorticity and buoyancy fields", flush=True)
    vor = np.zeros((3, n + 3, n + 3, n + 3))
    for c in range(3):
        g = W.gaussian_random_field(rng, n, n, n)   # (nx, ny, nz + 1) nodes, rms 0.1
        vor[c, 1:n + 1, 1:n + 1, 1:n + 2] = g
    x = lower[0] + (np.arange(n) + 0.0) * ext / n
    z = lower[2] + np.arange(n + 1) * ext / n
    X, Y, Z = np.meshgrid(x, x, z, indexing="ij", sparse=True)
    buoy = np.zeros((n + 3, n + 3, n + 3))
    amps = rng.normal(0.0, 0.5, 32)
    centres = lower[0] + rng.uniform(0.0, 1.0, (32, 3)) * ext
    for a, c0 in zip(amps, centres):
        buoy[1:n + 1, 1:n + 1, 1:n + 2] += a * np.exp(-((X - c0[0]) ** 2 + (Y - c0[1]) ** 2 + (Z - c0[2]) ** 2)
                                                       / (0.1 * ext) ** 2)
    from oracle import oracle as o

    d = o.make_dom(n, n, n, lower, extent)
    st = o.State(d, 0)
    st.vortg[:] = vor
    st.tbuoyg[:] = buoy
    gm = Member(n, n, n, lower, extent, max_parcels=1)
    gm.set_field("vortg", vor)
    gm.set_field("tbuoyg", buoy)
    del vor, buoy
    st.vor2vel()
    gm.vor2vel()
    for f in ("vortg", "velog", "velgradg"):
        e = maxrel(interior(gm.get_field(f), n, n), interior(getattr(st, f), n, n))
        _report("256^3 vor2vel", f, e)
        assert e <= TOL, (f, e)
    st.vorticity_tendency()
    gm.vorticity_tendency()
    e = maxrel(interior(gm.get_field("vtend"), n, n), interior(st.vtend, n, n))
    _report("256^3 vorticity_tendency", "vtend", e)
    assert e <= TOL, e
    gm.close()
