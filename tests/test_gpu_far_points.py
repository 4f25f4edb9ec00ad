"""Strongly stretched parcels whose support points lie beyond the
neighbour cells of the parcel's cell (several cells away in x/y).

The reference's par2grid and grid2par (parcel_interpl.f90:203-246,
371-431) take any support point whose 2x2x2 stencil lies in the grid with
its one-cell halo; the tile kernels handle points within one cell of the
parcel's cell from their LDS box and the rest through the member's global
fixed-point accumulator (par2grid) and a global gather (grid2par in the
push). Checked against the oracle to the parity tolerance, bitwise against
another parcel order, and, for points outside the x-y domain, against the
same parcels moved by whole cells."""
import numpy as np
import pytest

from epic_amd import workloads as W
from epic_amd.core import Batch, Member
from tests._util import maxrel, oracle_state

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _stretched(seed, count, a_cells, inside=True):
    """rising_thermal(16) with `count` parcels stretched in the x-y plane to
    a semi-major axis of `a_cells` cells (same volume, random orientation,
    the two minor axes 1.25 and 0.8 times their common mean),
    and the number of their support points beyond the neighbour cells.
    inside: only parcels whose support points stay inside the x-y domain
    (the reference's interpolation indexes a one-cell halo box without a
    periodic wrap, interpl.f90:17-56, so farther points are not valid input
    there)."""
    from oracle import oracle as o

    w = W.rising_thermal(16, seed=seed)
    rng = np.random.default_rng(seed)
    dx, lo, ext = np.asarray(w.dx), np.asarray(w.lower), np.asarray(w.extent)
    nfar = done = 0
    for i in rng.permutation(w.n):
        if done == count:
            break
        a = a_cells * dx[0]
        b = 1.25 * np.sqrt(3.0 * w.volume[i] / (4.0 * np.pi * a))   # c = 0.8 b from the volume: no equal axes
        th = rng.uniform(0.0, np.pi)
        c, s = np.cos(th), np.sin(th)
        Bi = [a * a * c * c + b * b * s * s, (a * a - b * b) * c * s, 0.0, a * a * s * s + b * b * c * c, 0.0]
        _, _, pts = o.get_points(w.position[i], Bi, w.volume[i])
        if inside and ((pts[:, :2] < lo[:2]).any() or (pts[:, :2] >= lo[:2] + ext[:2]).any()):
            continue
        w.B[i] = Bi
        cell = np.floor((w.position[i] - lo) / dx)
        nfar += int((np.abs(np.floor((pts - lo) / dx) - cell).max(axis=1) > 1).sum())
        done += 1
    return w, nfar


def _interior(a, w):
    return a[..., 1:w.nx + 1, 1:w.ny + 1, :]


@pytest.mark.parametrize("a_cells", [3.5, 6.0])
def test_par2grid_far_points(a_cells):
    w, nfar = _stretched(9, 64, a_cells)
    print(f"far points: {nfar} (semi-major axis {a_cells} cells)", flush=True)
    assert nfar > 40
    st = oracle_state(w)
    st.par2grid()
    gm = Member.from_workload(w)
    gm.par2grid()
    for f in ("volg", "vortg", "tbuoyg", "dbuoyg", "humg"):
        e = maxrel(_interior(gm.get_field(f), w), _interior(getattr(st, f), w))
        assert e <= TOL, (f, e)
    for f in ("nparg", "nsparg"):
        assert np.array_equal(_interior(gm.get_field(f), w), _interior(getattr(st, f), w)), f
    # the far accumulator starts from zero again: a second par2grid of the
    # same parcels gives the same bits
    first = gm.get_field("vortg")
    gm.par2grid()
    assert np.array_equal(gm.get_field("vortg"), first)


def test_par2grid_far_points_order_independent():
    w, _ = _stretched(10, 64, 3.5)
    a = Member.from_workload(w)
    a.par2grid()
    perm = np.random.default_rng(2).permutation(w.n)
    b = Member.from_workload(w)
    b.set_parcels(w.position[perm], w.vorticity[perm], w.B[perm], w.volume[perm], w.buoyancy[perm], w.humidity[perm])
    b.par2grid()
    for f in ("volg", "vortg", "tbuoyg", "dbuoyg", "humg", "nparg", "nsparg"):
        assert np.array_equal(a.get_field(f), b.get_field(f)), f


def test_grid2par_far_points():
    w, _ = _stretched(11, 64, 3.5)
    st = oracle_state(w)
    st.par2grid()
    st.vor2vel()
    st.vorticity_tendency()
    gm = Member.from_workload(w)
    for f in ("velog", "velgradg", "vtend"):
        gm.set_field(f, getattr(st, f))
    st.grid2par()
    gm.grid2par()
    dp, dv, s = gm.get_tendencies()
    for got, ref, f in ((dp, st.delta_pos, "delta_pos"), (dv, st.delta_vor, "delta_vor"), (s, st.strain, "strain")):
        assert maxrel(got, ref) <= TOL, (f, maxrel(got, ref))


def test_ls_rk_step_far_points():
    w, _ = _stretched(12, 64, 3.5)
    st = oracle_state(w)
    gm = Member.from_workload(w)
    t_ref = t_gpu = 0.0
    for _ in range(2):
        t_ref = st.ls_rk_step(t_ref, 0.01)
        t_gpu = gm.ls_rk_step(t_gpu, 0.01)
    got = gm.get_parcels()
    e = np.abs(got["position"] - st.position).max() / max(w.dx)
    assert e <= TOL, ("position", e)
    for f in ("vorticity", "B"):
        assert maxrel(got[f], getattr(st, f)) <= TOL, (f, maxrel(got[f], getattr(st, f)))
    for f in ("velog", "vtend", "vortg"):
        e = maxrel(_interior(gm.get_field(f), w), _interior(getattr(st, f), w))
        assert e <= TOL, (f, e)


def test_batch_with_far_points_equals_alone():
    """Each member has its own far accumulator: a batch of a stretched and
    a plain member steps bitwise like the two alone."""
    ws = [_stretched(13, 64, 3.5)[0], W.rising_thermal(16, seed=14)]
    alone = [Member.from_workload(w) for w in ws]
    inb = [Member.from_workload(w) for w in ws]
    b = Batch(inb)
    t = np.zeros(2)
    ta = [0.0, 0.0]
    for _ in range(2):
        for i, m in enumerate(alone):
            ta[i] = m.ls_rk_step(ta[i], 0.01)
        t, stat = b.step(t, np.array([0.01, 0.01]))
        assert (stat == 0).all()
    for x, y in zip(alone, inb):
        px, py = x.get_parcels(), y.get_parcels()
        for k in ("position", "vorticity", "B"):
            assert np.array_equal(px[k], py[k]), k
    b.close()


def test_far_points_across_the_periodic_edges():
    """Stretched parcels anywhere, support points up to three cells outside
    the x-y domain (no valid input for the reference): the grid is periodic
    in x and y, so moving every parcel by whole cells moves the gridded
    fields by as many cells and the step's parcels by the same distance."""
    from oracle import oracle as o

    w, _ = _stretched(15, 96, 3.5, inside=False)
    lo, ext, dx = np.asarray(w.lower), np.asarray(w.extent), np.asarray(w.dx)
    outside = 0
    for i in range(w.n):
        _, _, pts = o.get_points(w.position[i], w.B[i], w.volume[i])
        outside += int(((pts[:, :2] < lo[:2]) | (pts[:, :2] >= lo[:2] + ext[:2])).any(axis=1).sum())
    print(f"support points outside the x-y domain: {outside}", flush=True)
    assert outside > 10
    kx, ky = 5, 11
    pos = w.position.copy()
    pos[:, 0] += kx * dx[0]
    pos[:, 1] += ky * dx[1]
    for d in (0, 1):
        pos[:, d] = np.where(pos[:, d] >= lo[d] + ext[d], pos[:, d] - ext[d], pos[:, d])
    a = Member.from_workload(w)
    b = Member.from_workload(w)
    b.set_parcels(pos, w.vorticity, w.B, w.volume, w.buoyancy, w.humidity)
    a.par2grid()
    b.par2grid()
    for f in ("volg", "vortg", "tbuoyg", "nparg"):
        ga = np.roll(_interior(a.get_field(f), w), (kx, ky), axis=(-3, -2))
        gb = _interior(b.get_field(f), w)
        assert maxrel(gb, ga) <= TOL, (f, maxrel(gb, ga))
    ta = a.ls_rk_step(0.0, 0.01)
    tb = b.ls_rk_step(0.0, 0.01)
    assert ta == tb
    pa, pb = a.get_parcels(), b.get_parcels()
    shift = np.array([kx * dx[0], ky * dx[1], 0.0])
    dpos = (pb["position"] - pa["position"] - shift) / dx
    dpos[:, :2] -= np.round(dpos[:, :2] * dx[:2] / ext[:2]) * ext[:2] / dx[:2]   # whole periods
    assert np.abs(dpos).max() <= TOL, np.abs(dpos).max()
    assert maxrel(pb["vorticity"], pa["vorticity"]) <= TOL
    assert maxrel(pb["B"], pa["B"]) <= TOL
