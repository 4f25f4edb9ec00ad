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


_MAKERS = {
    "beltrami16": lambda: W.beltrami(16),
    "thermal16": lambda: W.rising_thermal(16, seed=3),
    "moist16": lambda: W.rising_thermal(16, seed=4, humidity_plume=True, n_blobs=4),
    "thermal32x16x24": lambda: W.rising_thermal(32, ny=16, nz=24, seed=5),
    # radix-3 lengths in y and in the z transforms (ny = 24, 2 nz = 24)
    "thermal16x24x12": lambda: W.rising_thermal(16, ny=24, nz=12, seed=6),
    # the batched ensemble's radix-3 lengths in x (the x pass transforms
    # nx/2 = 24 and 48 complex points) and y
    "thermal48x32x32": lambda: W.rising_thermal(48, ny=32, nz=32, seed=7),
    "thermal96x48x32": lambda: W.rising_thermal(96, ny=48, nz=32, seed=8),
    # radix-5 lengths in every transform (nx/2 = 20, ny = 40, 2 nz = 40)
    "thermal40x40x20": lambda: W.rising_thermal(40, ny=40, nz=20, seed=9),
    # partial tiles (nx, ny not multiples of 8: 8 + 8 + 4 and 4 x 8 + 4
    # cells), radix 5 in x (nx/2 = 10) and z (2 nz = 20), radix 3 in y
    "thermal20x36x10": lambda: W.rising_thermal(20, ny=36, nz=10, seed=10),
    # fewer than 11 nodes across (a tile's node box wraps onto itself) and
    # partial tiles in both directions
    "thermal10x12x8": lambda: W.rising_thermal(10, ny=12, nz=8, seed=11),
}


class _Members(dict):
    """Workloads generated on first use (not at collection)."""

    def __missing__(self, name):
        mem = _MAKERS[name]()
        self[name] = mem
        return mem

    def __iter__(self):
        return iter(_MAKERS)


MEMBERS = _Members()


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
    # integer sums per tile and a fixed order of the tile sums: the same bits
    for f in ("volg", "vortg", "tbuoyg", "dbuoyg", "humg", "nparg", "nsparg"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f
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
@pytest.mark.parametrize("name,nsteps", [("beltrami16", 2), ("thermal16", 4), ("moist16", 3), ("thermal32x16x24", 1),
                                          ("thermal48x32x32", 2), ("thermal96x48x32", 1), ("thermal40x40x20", 2),
                                          ("thermal20x36x10", 2), ("thermal10x12x8", 2)])
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
    for f in ("velog", "velgradg", "vtend", "volg", "vortg", "tbuoyg"):
        e = maxrel(interior(gm.get_field(f), nx, ny, 0, None), interior(getattr(st, f), nx, ny, 0, None))
        print(f"  {name} x{nsteps} {f}: {e:.2e}", flush=True)
        assert e <= TOL, (name, f, e)


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


def test_ls_rk_step_reflective_walls():
    """Parcels whose centres lie outside the z walls (within the halo
    layer) stay outside during the step (w vanishes at the walls) and are
    mirrored by apply_reflective_bc after the last stage
    (parcel_bc.f90:26-40, ls_rk.f90:126): z -> 2 wall - z, B13 and B23
    change sign."""
    mem = W.rising_thermal(16, seed=12)
    rng = np.random.default_rng(5)
    dz = mem.dx[2]
    lo, hi = mem.lower[2], mem.lower[2] + mem.extent[2]
    pos = mem.position.copy()
    bot = np.flatnonzero(pos[:, 2] < lo + 0.5 * dz)[:200]
    top = np.flatnonzero(pos[:, 2] > hi - 0.5 * dz)[:200]
    pos[bot, 2] = lo - rng.uniform(0.05, 0.3, bot.size) * dz
    pos[top, 2] = hi + rng.uniform(0.05, 0.3, top.size) * dz
    mem.position = pos
    st = oracle_state(mem)
    gm = Member.from_workload(mem)
    st.ls_rk_step(0.0, 0.01)
    gm.ls_rk_step(0.0, 0.01)
    # the oracle mirrored them: every centre is inside, B13/B23 flipped
    assert (st.position[:, 2] >= lo).all() and (st.position[:, 2] <= hi).all()
    assert np.abs(st.position[bot, 2] - (2 * lo - pos[bot, 2])).max() < 0.05 * dz
    assert np.abs(st.position[top, 2] - (2 * hi - pos[top, 2])).max() < 0.05 * dz
    assert (np.sign(st.B[bot, 2]) == -np.sign(mem.B[bot, 2])).mean() > 0.9
    got = gm.get_parcels()
    e = np.abs(got["position"] - st.position).max() / max(mem.dx)
    assert e <= TOL, ("position", e)
    for f in ("vorticity", "B"):
        assert maxrel(got[f], getattr(st, f)) <= TOL, (f, maxrel(got[f], getattr(st, f)))
    moved = np.concatenate([bot, top])
    assert np.abs(got["position"][moved] - st.position[moved]).max() <= TOL * max(mem.dx)
    assert maxrel(got["B"][moved], st.B[moved]) <= TOL


def _zero_mode_and_filter_vortg(nx, ny, nz, lower, extent):
    """Horizontally uniform xi(z), eta(z) (the mean flow of
    inversion.f90:161-176) plus a high horizontal mode of zeta with a high
    sine mode in z (attenuated by the Hou-Li filter)."""
    x = lower[0] + (np.arange(nx + 3) - 1) * extent[0] / nx
    y = lower[1] + (np.arange(ny + 3) - 1) * extent[1] / ny
    z = (np.arange(nz + 3) - 1) * extent[2] / nz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    Lz = extent[2]
    vor = np.zeros((3,) + X.shape)
    vor[0] = -0.2 + 0.9 * Z / Lz - 0.6 * np.sin(2 * np.pi * Z / Lz)
    vor[1] = 0.7 - 1.3 * Z / Lz + 0.45 * np.sin(3 * np.pi * Z / Lz)
    vor[2] = np.cos(2 * np.pi * (nx // 2 - 1) * (X - lower[0]) / extent[0]) * np.sin((nz - 1) * np.pi * Z / Lz)
    return vor


def test_vor2vel_mean_flow_and_filter():
    """GPU vs oracle on the zero mode and the filter, which random fields
    exercise only weakly."""
    nx, ny, nz = 32, 16, 24
    lower, extent = (-1.0, -2.0, -0.3), (2.0, 3.0, 1.7)
    from oracle import oracle as o

    d = o.make_dom(nx, ny, nz, lower, extent)
    st = o.State(d, 0)
    vor = _zero_mode_and_filter_vortg(nx, ny, nz, lower, extent)
    st.vortg[:] = vor
    st.vor2vel()
    gm = Member(nx, ny, nz, lower, extent, max_parcels=1)
    gm.set_field("vortg", vor)
    gm.vor2vel()
    for f in ("vortg", "velog", "velgradg"):
        e = maxrel(interior(gm.get_field(f), nx, ny, 1, nz + 2), interior(getattr(st, f), nx, ny, 1, nz + 2))
        print(f"  mean flow + filter {f}: {e:.2e}", flush=True)
        assert e <= TOL, (f, e)


def test_create_rejects_grid_beyond_transform_lds():
    """16 lines of a 320-point FP64 y transform need 169 KB of LDS, more than
    a work-group may hold: epic_create refuses the grid (EPIC_E_INVALID)
    instead of a later launch failing."""
    with pytest.raises(EpicError) as ei:
        Member(16, 320, 16, (0.0, 0.0, 0.0), (1.0, 20.0, 1.0))
    assert ei.value.code == -1
    m = Member(16, 300, 16, (0.0, 0.0, 0.0), (1.0, 20.0, 1.0))   # 155 KB: accepted
    m.close()
