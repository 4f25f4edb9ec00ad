"""GPU parity against the FP64 oracle, through the C-ABI.

Tolerances: gridded fields and parcel attributes max|x - x_ref| / max|x_ref|
<= 1e-5 (FP32 storage, FP64 positions/eigen/z-transforms); integer work
(parcel counts, cell sort) bit-exact."""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import EpicError, Member
from tests._util import maxrel, oracle_state

pytestmark = pytest.mark.gpu
TOL = 1e-5


def interior(a, nx, ny, z0=1, z1=None):
    """x/y interior, z planes [z0-1 .. z1-1] (array index = iz + 1)."""
    z1 = a.shape[-1] - 1 if z1 is None else z1
    return a[..., 1:nx + 1, 1:ny + 1, z0:z1]


def _members():
    return {
        "beltrami16": W.beltrami(16),
        "thermal16": W.rising_thermal(16, seed=3),
        "moist16": W.rising_thermal(16, seed=4, humidity_plume=True, n_blobs=4),
        "thermal32x16x24": W.rising_thermal(32, ny=16, nz=24, seed=5),
        # radix-3 lengths in y and in the z transforms (ny = 24, 2 nz = 24)
        "thermal16x24x12": W.rising_thermal(16, ny=24, nz=12, seed=6),
    }


MEMBERS = _members()


@pytest.mark.parametrize("name", list(MEMBERS))
def test_par2grid(name):
    mem = MEMBERS[name]
    st = oracle_state(mem)
    st.par2grid()
    gm = Member.from_workload(mem)
    gm.par2grid()
    nx, ny = mem.nx, mem.ny
    for f in ("volg", "vortg", "tbuoyg", "dbuoyg", "humg"):
        ref = getattr(st, f)
        got = gm.get_field(f)
        # all z planes -1..nz+1 (halo planes carry the reference's raw / extrapolated values)
        e = maxrel(interior(got, nx, ny, 0, None), interior(ref, nx, ny, 0, None))
        assert e <= TOL, (name, f, e)
    for f in ("nparg", "nsparg"):
        assert np.array_equal(interior(gm.get_field(f), nx, ny, 0, None), interior(getattr(st, f), nx, ny, 0, None)), f


def test_par2grid_deterministic():
    mem = MEMBERS["thermal32x16x24"]
    a = Member.from_workload(mem)
    a.par2grid()
    b = Member.from_workload(mem)
    # same parcel set handed over in another order: identical bits
    perm = np.random.default_rng(1).permutation(mem.n)
    b.set_parcels(mem.position[perm], mem.vorticity[perm], mem.B[perm], mem.volume[perm], mem.buoyancy[perm],
                  mem.humidity[perm])
    b.par2grid()
    a.par2grid()
    for f in ("volg", "vortg", "tbuoyg"):
        fa, fb = a.get_field(f), b.get_field(f)
        assert np.array_equal(fa, fb) or maxrel(fb, fa) <= TOL, f
    c = Member.from_workload(mem)
    c.par2grid()
    for f in ("volg", "vortg", "tbuoyg", "dbuoyg", "humg"):
        assert np.array_equal(a.get_field(f), c.get_field(f)), f


def test_par2grid_scales_follow_the_current_parcels():
    """The fixed-point scales come from the parcels of this par2grid: after
    one with vorticity a million times larger, the same grid is within the
    tolerance (left-over maxima would coarsen it by that factor)."""
    mem = MEMBERS["thermal16"]
    gm = Member.from_workload(mem)
    gm.set_parcels(mem.position, mem.vorticity * 1e6, mem.B, mem.volume, mem.buoyancy, mem.humidity)
    gm.par2grid()
    gm.set_parcels(mem.position, mem.vorticity, mem.B, mem.volume, mem.buoyancy, mem.humidity)
    gm.par2grid()
    st = oracle_state(mem)
    st.par2grid()
    nx, ny = mem.nx, mem.ny
    e = maxrel(interior(gm.get_field("vortg"), nx, ny, 0, None), interior(st.vortg, nx, ny, 0, None))
    assert e <= TOL, e


def _taylor_green(nx, ny, nz):
    lower = (-np.pi, -np.pi, -0.5 * np.pi)
    extent = (2 * np.pi,) * 3
    dx = [extent[i] / (nx, ny, nz)[i] for i in range(3)]
    x = lower[0] + (np.arange(nx + 3) - 1) * dx[0]
    y = lower[1] + (np.arange(ny + 3) - 1) * dx[1]
    z = lower[2] + (np.arange(nz + 3) - 1) * dx[2]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    A, B, Cc = 1.0, 1.0, -2.0
    vor = np.stack([(Cc - B) * np.sin(X) * np.cos(Y) * np.cos(Z), (A - Cc) * np.cos(X) * np.sin(Y) * np.cos(Z),
                    (B - A) * np.cos(X) * np.cos(Y) * np.sin(Z)])
    vel = np.stack([A * np.cos(X) * np.sin(Y) * np.sin(Z), B * np.sin(X) * np.cos(Y) * np.sin(Z),
                    Cc * np.sin(X) * np.sin(Y) * np.cos(Z)])
    return lower, extent, vor, vel


def test_vor2vel_taylor_green():
    """test_mpi_vor2vel.f90 closed form on the GPU (FP32 fields)."""
    nx = ny = nz = 32
    lower, extent, vor, vel = _taylor_green(nx, ny, nz)
    gm = Member(nx, ny, nz, lower, extent, max_parcels=1)
    gm.set_field("vortg", vor)
    gm.vor2vel()
    got = gm.get_field("velog")
    e = maxrel(interior(got, nx, ny), interior(vel, nx, ny))
    assert e < 2e-6, e


@pytest.mark.parametrize("name", list(MEMBERS))
def test_vor2vel_and_vtend_same_input(name):
    """Feed the oracle's par2grid fields to both inversions."""
    mem = MEMBERS[name]
    st = oracle_state(mem)
    st.par2grid()
    gm = Member(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, mem.glat, mem.lambda_c, mem.q0, mem.f_cor,
                mem.zeta_zero, max_parcels=1)
    gm.set_field("vortg", st.vortg)
    gm.set_field("tbuoyg", st.tbuoyg)
    st.vor2vel()
    gm.vor2vel()
    nx, ny = mem.nx, mem.ny
    for f in ("vortg", "velog", "velgradg"):
        e = maxrel(interior(gm.get_field(f), nx, ny, 0, None), interior(getattr(st, f), nx, ny, 0, None))
        assert e <= TOL, (name, f, e)
    st.vorticity_tendency()
    gm.vorticity_tendency()
    e = maxrel(interior(gm.get_field("vtend"), nx, ny, 0, None), interior(st.vtend, nx, ny, 0, None))
    assert e <= TOL, (name, "vtend", e)


@pytest.mark.parametrize("name", list(MEMBERS))
def test_grid2par_same_fields(name):
    mem = MEMBERS[name]
    st = oracle_state(mem)
    st.par2grid()
    st.vor2vel()
    st.vorticity_tendency()
    gm = Member.from_workload(mem)
    for f in ("velog", "velgradg", "vtend"):
        gm.set_field(f, getattr(st, f))
    st.grid2par()
    gm.grid2par()
    dp, dv, s = gm.get_tendencies()
    for got, ref, f in ((dp, st.delta_pos, "delta_pos"), (dv, st.delta_vor, "delta_vor"), (s, st.strain, "strain")):
        assert maxrel(got, ref) <= TOL, (name, f, maxrel(got, ref))


def test_grid2par_constant_fields():
    """test_mpi_grid2par.f90: constant fields come back exactly (FP32)."""
    mem = W.rising_thermal(16, seed=2)
    gm = Member.from_workload(mem)
    H = (mem.nx + 3, mem.ny + 3, mem.nz + 3)
    gm.set_field("velog", np.stack([np.full(H, v) for v in (1.0, 2.0, 3.0)]))
    gm.set_field("velgradg", np.stack([np.full(H, float(v)) for v in range(1, 9)]))
    gm.set_field("vtend", np.stack([np.full(H, v) for v in (10.0, 20.0, 30.0)]))
    gm.grid2par()
    dp, dv, s = gm.get_tendencies()
    assert np.abs(dp - [1.0, 2.0, 3.0]).max() <= 4e-7 * 3
    assert np.abs(s - np.arange(1, 9)).max() <= 4e-7 * 8
    assert np.abs(dv - [10.0, 20.0, 30.0]).max() <= 4e-7 * 30


# thermal16 over 4 steps includes a physical re-sort (every 3 steps) and the
# fixed-point maxima stored by the pushes; moist16 over 3 steps, their
# humidity bounds
@pytest.mark.parametrize("name,nsteps", [("beltrami16", 2), ("thermal16", 4), ("moist16", 3), ("thermal32x16x24", 1)])
def test_ls_rk_step(name, nsteps):
    mem = MEMBERS[name]
    st = oracle_state(mem)
    gm = Member.from_workload(mem)
    t_ref = t_gpu = 0.0
    for _ in range(nsteps):
        t_ref = st.ls_rk_step(t_ref, 0.01)
        t_gpu = gm.ls_rk_step(t_gpu, 0.01)
    assert t_ref == t_gpu
    got = gm.get_parcels()
    for f in ("position", "vorticity", "B"):
        ref = getattr(st, f)
        if f == "position":
            e = np.abs(got[f] - ref).max() / max(mem.dx)   # positions: relative to the cell size
        else:
            e = maxrel(got[f], ref)
        print(f"  {name} x{nsteps} {f}: {e:.2e}", flush=True)
        assert e <= TOL, (name, f, e)
    nx, ny = mem.nx, mem.ny
    for f in ("velog", "vtend", "volg", "vortg"):
        e = maxrel(interior(gm.get_field(f), nx, ny, 0, None), interior(getattr(st, f), nx, ny, 0, None))
        assert e <= 3 * TOL, (name, f, e)


def test_ls_rk_step_bitwise_reproducible():
    mem = MEMBERS["thermal16"]
    runs = []
    for _ in range(2):
        gm = Member.from_workload(mem)
        t = 0.0
        for _ in range(2):
            t = gm.ls_rk_step(t, 0.01)
        runs.append(gm.get_parcels())
    for k in runs[0]:
        assert np.array_equal(runs[0][k], runs[1][k]), k


def test_cell_sort_bit_exact():
    """parcel_nearest.f90 counting sort: loca, nppc, kc1 and node bit-exact."""
    from oracle import oracle as o

    mem = W.rising_thermal(16, ny=24, nz=8, seed=8)
    rng = np.random.default_rng(4)
    perm = rng.permutation(mem.n)
    pos = mem.position[perm].copy()
    dx = np.asarray(mem.dx)
    lo = np.asarray(mem.lower)
    # edge cases: exactly on cell faces, on the lower periodic edge, on the z walls
    pos[0] = lo
    pos[1] = lo + np.asarray(mem.extent) - 1e-12
    pos[2, 0] = lo[0] + 3 * dx[0]
    pos[3, 2] = lo[2] + mem.extent[2]
    gm = Member(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, max_parcels=mem.n)
    gm.set_parcels(pos, mem.vorticity, mem.B, mem.volume, mem.buoyancy, mem.humidity)
    d = o.make_dom(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent)
    ref = o.nearest_cell_sort(d, pos)
    got = gm.cell_sort()
    for a, b, name in zip(got, ref, ("loca", "nppc", "kc1", "node")):
        assert np.array_equal(a, b), name


def test_errors_are_reported():
    mem = W.rising_thermal(16, seed=6)
    pos = mem.position.copy()
    pos[5, 2] = mem.lower[2] - 2.5 * mem.dx[2]   # two cells below the wall
    gm = Member(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, max_parcels=mem.n)
    gm.set_parcels(pos, mem.vorticity, mem.B, mem.volume, mem.buoyancy, mem.humidity)
    with pytest.raises(EpicError) as e:
        gm.par2grid()
    assert e.value.code == -3
    gm2 = Member.from_workload(mem)
    with pytest.raises(EpicError) as e:
        gm2.ls_rk_step(0.0, 0.0)
    assert e.value.code == -7
    B = mem.B.copy()
    B[7] = [1e-9, 1e-9, 0.0, 1e-9, 0.0]   # B11 B22 - B12^2 = 0
    gm3 = Member(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent, max_parcels=mem.n)
    gm3.set_parcels(mem.position, mem.vorticity, B, mem.volume, mem.buoyancy, mem.humidity)
    with pytest.raises(EpicError) as e:
        gm3.par2grid()
    assert e.value.code == -4
