"""Oracle par2grid / grid2par / ls_rk_step / cell sort against the reference's
known-answer tests, invariants and brute force."""
import math

import numpy as np
import pytest

from epic_amd import workloads as W
from tests._util import oracle_state


def _lattice_state(orc, n, ext, lower, n_per_dim):
    d = orc.make_dom(n, n, n, (lower,) * 3, (ext,) * 3)
    dx = (ext / n,) * 3
    pos = W.lattice_positions(n, n, n, (lower,) * 3, dx, n_per_dim)
    st = orc.State(d, pos.shape[0])
    st.position[:] = pos
    vol = d.vcell / n_per_dim ** 3
    st.volume[:] = vol
    l23 = (0.75 * vol / math.pi) ** (2.0 / 3.0)
    st.B[:, 0] = l23
    st.B[:, 3] = l23
    return st


def test_trilinear_sum_rules(orc):
    """test_mpi_trilinear.f90: 32^3, 27 parcels/cell, omega_x = 1.5:
    sum(vortg_x) = 1.5 ngrid and sum(volg) = ngrid vcell (:53-59)."""
    st = _lattice_state(orc, 32, 0.4, -1.5, 3)
    st.vorticity[:, 0] = 1.5
    st.par2grid()
    d = st.dom
    ngrid = d.nx * d.ny * (d.nz + 1)
    interior = (slice(1, d.nx + 1), slice(1, d.ny + 1), slice(1, d.nz + 2))
    assert abs(st.vortg[0][interior].sum() - 1.5 * ngrid) < 1e-9
    assert abs(st.volg[interior].sum() - ngrid * d.vcell) < 3e-14 * ngrid
    assert st.nparg[interior].sum() == st.n
    # every node sees exactly one cell volume after the free-slip folds
    assert np.abs(st.volg[interior] / d.vcell - 1.0).max() < 1e-12


def test_grid2par_constant_fields(orc):
    """test_mpi_grid2par.f90: constant fields are interpolated exactly (1e-15)."""
    st = _lattice_state(orc, 32, 2.0, -1.0, 3)
    for l in range(3):
        st.velog[l][:] = l + 1.0
        st.vtend[l][:] = 10.0 * (l + 1)
    for l in range(8):
        st.velgradg[l][:] = l + 1.0
    st.grid2par()
    for l in range(3):
        assert np.abs(st.delta_pos[:, l] - (l + 1.0)).max() < 1e-15
        assert np.abs(st.delta_vor[:, l] - 10.0 * (l + 1)).max() < 1e-13
    for l in range(5):   # the reference checks strain components 1..5 at 1e-15
        assert np.abs(st.strain[:, l] - (l + 1.0)).max() < 1e-15
    for l in range(5, 8):
        assert np.abs(st.strain[:, l] - (l + 1.0)).max() < 4 * np.spacing(l + 1.0)


def test_par2grid_brute_force(orc):
    """Small random ellipsoids: par2grid against a direct numpy restatement
    of the scatter + halo fold + free-slip + normalisation."""
    mem = W.rising_thermal(8, seed=5)
    st = oracle_state(mem)
    st.par2grid()
    d = st.dom
    nx, ny, nz = d.nx, d.ny, d.nz
    raw = np.zeros((7, nx, ny, nz + 3))   # z index iz+1
    for n in range(st.n):
        ve, vt, pts = orc.get_points(st.position[n], st.B[n], st.volume[n])
        z = st.position[n, 2]
        qc = max(0.0, st.humidity[n] - d.q0 * math.exp(d.lambda_c * (d.lower[2] - z)))
        vals = [1.0, *st.vorticity[n], st.buoyancy[n] + d.glat * qc, st.buoyancy[n], st.humidity[n]]
        for p in pts:
            xyz = (p - np.array(d.lower[:])) / np.array(d.dx[:])
            i0 = np.floor(xyz).astype(int)
            f = xyz - i0
            for kk in (0, 1):
                for jj in (0, 1):
                    for ii in (0, 1):
                        w = (f[0] if ii else 1 - f[0]) * (f[1] if jj else 1 - f[1]) * (f[2] if kk else 1 - f[2])
                        for q in range(7):
                            raw[q, (i0[0] + ii) % nx, (i0[1] + jj) % ny, i0[2] + kk + 1] += 0.25 * st.volume[n] * w * vals[q]
    raw[:, :, :, 1] *= 2
    raw[:, :, :, nz + 1] *= 2
    raw[:, :, :, 2] += raw[:, :, :, 0]
    raw[:, :, :, nz] += raw[:, :, :, nz + 2]
    vol = raw[0, :, :, 1:nz + 2]
    got = {"volg": st.volg, "tbuoyg": st.tbuoyg, "dbuoyg": st.dbuoyg, "humg": st.humg}
    sl = (slice(1, nx + 1), slice(1, ny + 1), slice(1, nz + 2))
    assert np.allclose(got["volg"][sl], vol, rtol=1e-13, atol=0)
    for q, name in ((4, "tbuoyg"), (5, "dbuoyg"), (6, "humg")):
        assert np.allclose(got[name][sl], raw[q, :, :, 1:nz + 2] / vol, rtol=1e-12, atol=1e-14)
    for c in range(3):
        assert np.allclose(st.vortg[c][sl], raw[1 + c, :, :, 1:nz + 2] / vol, rtol=1e-12, atol=1e-14)


def test_par2grid_parcel_count_error(orc):
    """A parcel two cells below the domain is lost from the counts:
    "Wrong total number of parcels" (parcel_interpl.f90:248-250)."""
    st = _lattice_state(orc, 8, 1.0, 0.0, 2)
    st.position[0, 2] = -0.15   # 1.2 cells below lower
    with pytest.raises(orc.OracleError) as e:
        st.par2grid()
    assert e.value.code in (-1, -3)


def test_nearest_cell_sort_brute_force(orc):
    """Counting sort of parcel_nearest.f90:196-257: stable in parcel order,
    key ijk = 1 + ix + hx iy + hx hy iz on the halo-inclusive box."""
    mem = W.rising_thermal(8, seed=9)
    d = orc.make_dom(mem.nx, mem.ny, mem.nz, mem.lower, mem.extent)
    rng = np.random.default_rng(2)
    pos = mem.position[rng.permutation(mem.n)]
    loca, nppc, kc1, node = orc.nearest_cell_sort(d, pos)
    hx, hy = d.nx + 3, d.ny + 3
    ix = np.trunc((pos[:, 0] - (d.lower[0] - d.dx[0])) * d.dxi[0]).astype(int)
    iy = np.trunc((pos[:, 1] - (d.lower[1] - d.dx[1])) * d.dxi[1]).astype(int)
    iz = np.minimum(np.trunc((pos[:, 2] - d.lower[2]) * d.dxi[2]).astype(int), d.nz - 1)
    key = 1 + ix + hx * iy + hx * hy * iz
    assert np.array_equal(loca, key)
    assert np.array_equal(node, np.argsort(key, kind="stable"))
    assert kc1[0] == 1 and np.array_equal(np.diff(kc1), nppc[:-1])
    assert len(kc1) == hx * hy * (d.nz + 2)


def test_periodic_bc_truncation(orc):
    """apply_periodic_bc uses int() truncation (parcel_bc.f90:17-21): a parcel
    exactly at lower maps to upper."""
    d = orc.make_dom(8, 8, 8, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    import ctypes as C
    for x, want in ((0.0, 1.0), (0.3, 0.3), (1.0, 0.0), (1.2, 0.2), (-0.2, 0.8)):
        p = np.array([x, 0.5, 0.5])
        orc.lib().orc_apply_periodic_bc(C.byref(d), orc.ptr(p))
        assert abs(p[0] - want) < 1e-15, (x, p[0])


@pytest.mark.parametrize("order,p", [(4, 4), (3, 3)])
def test_lsrk_stability_polynomial(orc, order, p):
    """The low-storage coefficients (ls_rk.f90:31-54) make one step of
    y' = lam y reproduce exp(lam) through lam^p: the error ratio on halving
    lam is 2^(p+1) (order conditions for linear problems)."""
    L = orc.lib()
    errs = [abs(L.orc_lsrk_linear(z, order) - math.exp(z)) for z in (0.2, 0.1, 0.05)]
    for a, b in zip(errs, errs[1:]):
        assert 0.8 * 2 ** (p + 1) < a / b < 1.25 * 2 ** (p + 1), errs
    # 5-stage scheme: R(z) = sum_{k<=4} z^k/k! + c5 z^5 with c5 != 1/120 in general
    z = 1e-2
    taylor = sum(z ** k / math.factorial(k) for k in range(p + 1))
    assert abs(L.orc_lsrk_linear(z, order) - taylor) < 2 * z ** (p + 1)


def test_rk_step_volume_and_beltrami_steady(orc):
    """One RK step of the Beltrami member: B33 re-derived keeps det B =
    (3V/4pi)^2, and omega = alpha u stays (steady Beltrami flow: the
    vorticity tendency nearly vanishes)."""
    mem = W.beltrami(16)
    st = oracle_state(mem)
    st.ls_rk_step(0.0, 0.01)
    dv = np.abs(st.vorticity - mem.vorticity).max()
    assert dv < 1e-2 * np.abs(mem.vorticity).max(), dv
    for n in range(0, st.n, 997):
        b33 = orc.get_B33(st.B[n], st.volume[n])
        b = st.B[n]
        det = b[0] * (b[3] * b33 - b[4] ** 2) - b[1] * (b[1] * b33 - b[2] * b[4]) + b[2] * (b[1] * b[4] - b[2] * b[3])
        assert abs(det - (0.75 * st.volume[n] / math.pi) ** 2) < 1e-12 * det
