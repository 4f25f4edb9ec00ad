"""Oracle functions pinned by hand-computed cases and closed forms that a
sign, index or ordering mistake fails:

- apply_reflective_bc (parcel_bc.f90:26-40): mirror about the wall that was
  crossed, sign flip of B13 and B23 only, strict inequalities;
- the horizontally-averaged flow of vor2vel (inversion.f90:161-176, with
  gamtop/gambot of inversion_utils.f90:222-238): for vorticity
  (xi(z), eta(z), 0) the velocity is (ubar, vbar, 0) with
  d ubar/dz = eta, d vbar/dz = -xi and zero vertical mean;
- the Hou-Li filter (inversion_utils.f90:130-150): a pure mode
  zeta = cos(k x) sin(m pi z'/Lz) leaves vor2vel multiplied by
  exp(-36 (k/kmax)^36 - 36 (m/nz)^36), with kmax = pi nx / Lx;
- grid2par (parcel_interpl.f90:405-470) on linear fields: trilinear
  interpolation is exact for them, and the four support points average to
  the centre, so each parcel receives the field's value at its centre.
"""
import ctypes as C
import math

import numpy as np
import pytest

from epic_amd import workloads as W
from tests._util import oracle_state

PI = math.pi


def _reflect(orc, d, z, B):
    p = np.array([0.1, 0.2, z])
    b = np.array(B, dtype=np.float64)
    orc.lib().orc_apply_reflective_bc(C.byref(d), orc.ptr(p), orc.ptr(b))
    return p, b


def test_reflective_bc_hand_cases(orc):
    """Walls at z = -0.5 and z = 1.5 (lower -0.5, extent 2)."""
    d = orc.make_dom(8, 8, 8, (0.0, 0.0, -0.5), (1.0, 1.0, 2.0))
    B = [1.0, 2.0, 3.0, 4.0, 5.0]
    # above the upper wall: z -> 2 upper - z = 3.0 - 1.6
    p, b = _reflect(orc, d, 1.6, B)
    assert p[2] == pytest.approx(1.4, abs=1e-15) and tuple(p[:2]) == (0.1, 0.2)
    assert list(b) == [1.0, 2.0, -3.0, 4.0, -5.0]
    # below the lower wall: z -> 2 lower - z = -1.0 - (-0.75)
    p, b = _reflect(orc, d, -0.75, B)
    assert p[2] == pytest.approx(-0.25, abs=1e-15)
    assert list(b) == [1.0, 2.0, -3.0, 4.0, -5.0]
    # exactly on either wall, and inside: unchanged (strict > and <)
    for z in (1.5, -0.5, 0.3):
        p, b = _reflect(orc, d, z, B)
        assert p[2] == z and list(b) == B


def _zero_mode_dom(orc):
    return orc.make_dom(8, 8, 16, (-1.0, -1.0, -0.3), (2.0, 2.0, 1.7))


def _zgrid(orc, d, fz):
    """Halo array with f(z) on the interior nodes of planes 0..nz."""
    nx, ny, nz = d.nx, d.ny, d.nz
    z = np.arange(nz + 1) * d.dx[2]   # z' = z - lower
    a = np.zeros((nx + 3, ny + 3, nz + 3))
    a[1:nx + 1, 1:ny + 1, 1:nz + 2] = fz(z)[None, None, :]
    return a


def test_vor2vel_horizontal_mean_flow_closed_form(orc):
    """xi(z) = a2 + b2 z'/Lz + c2 sin(2 pi z'/Lz), eta(z) = a + b z'/Lz +
    c sin(3 pi z'/Lz): ubar = a (z' - Lz/2) + b (z'^2/(2 Lz) - Lz/6)
    - c Lz/(3 pi) cos(3 pi z'/Lz), and vbar the same for -xi; w = 0."""
    d = _zero_mode_dom(orc)
    Lz = d.extent[2]
    st = orc.State(d, 0)
    a, b, c, m = 0.7, -1.3, 0.45, 3
    a2, b2, c2, m2 = -0.2, 0.9, -0.6, 2
    eta = lambda z: a + b * z / Lz + c * np.sin(m * PI * z / Lz)  # noqa: E731
    xi = lambda z: a2 + b2 * z / Lz + c2 * np.sin(m2 * PI * z / Lz)  # noqa: E731

    def integral(a, b, c, m, z):   # zero-mean antiderivative on [0, Lz]
        return a * (z - Lz / 2) + b * (z * z / (2 * Lz) - Lz / 6) - c * Lz / (m * PI) * np.cos(m * PI * z / Lz)

    st.vortg[0] = _zgrid(orc, d, xi)
    st.vortg[1] = _zgrid(orc, d, eta)
    st.vor2vel()
    sl = (slice(1, d.nx + 1), slice(1, d.ny + 1), slice(1, d.nz + 2))
    ubar = _zgrid(orc, d, lambda z: integral(a, b, c, m, z))
    vbar = _zgrid(orc, d, lambda z: -integral(a2, b2, c2, m2, z))
    assert np.abs(st.velog[0] - ubar)[sl].max() < 1e-13
    assert np.abs(st.velog[1] - vbar)[sl].max() < 1e-13
    assert np.abs(st.velog[2])[sl].max() < 1e-14
    # vortg keeps the horizontally uniform components (no filter on (0,0))
    assert np.abs(st.vortg[0] - _zgrid(orc, d, xi))[sl].max() < 1e-13
    assert np.abs(st.vortg[1] - _zgrid(orc, d, eta))[sl].max() < 1e-13
    # du/dz = eta and dv/dz = -xi through vel2vgrad's identities
    assert np.abs(st.velgradg[2] - _zgrid(orc, d, eta))[sl].max() < 1e-13
    assert np.abs(st.velgradg[5] + _zgrid(orc, d, xi))[sl].max() < 1e-13


@pytest.mark.parametrize("axis,k,m", [(0, 7, 15), (1, 6, 3), (0, 3, 14), (1, 8, 1)])
def test_vor2vel_hou_li_filter(orc, axis, k, m):
    """zeta = cos(k x) sin(m pi z'/Lz) (zero on the walls, so a pure sine
    mode with no harmonic part) comes out of vor2vel scaled by the filter."""
    nx, ny, nz = 16, 16, 16
    d = orc.make_dom(nx, ny, nz, (-1.0, -2.0, 0.0), (2.0, 3.0, 1.5))
    st = orc.State(d, 0)
    L = d.extent[axis]
    nh = (nx, ny)[axis]
    x = d.lower[0] + np.arange(nx) * d.dx[0]
    y = d.lower[1] + np.arange(ny) * d.dx[1]
    z = np.arange(nz + 1) * d.dx[2]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    H = (X, Y)[axis]
    zeta = np.cos(2 * PI * k * (H - d.lower[axis]) / L) * np.sin(m * PI * Z / d.extent[2])
    st.vortg[2][1:nx + 1, 1:ny + 1, 1:nz + 2] = zeta
    st.vor2vel()
    kmax = PI * nh / L
    kk = 2 * PI * k / L
    F = math.exp(-36.0 * (kk / kmax) ** 36 - 36.0 * (m / nz) ** 36)
    got = st.vortg[2][1:nx + 1, 1:ny + 1, 1:nz + 2]
    assert np.abs(got - F * zeta).max() < 1e-13, (F, np.abs(got - F * zeta).max())
    # the attenuation is visible at the high modes (the test would not
    # notice a missing filter otherwise)
    if k >= 7 or m >= 14:
        assert F < 0.8


def test_grid2par_linear_fields_exact(orc):
    """Random triaxial ellipsoids: every gathered component equals the
    linear field at the parcel centre (FP64 rounding only)."""
    mem = W.rising_thermal(8, seed=4, ny=12, nz=10)
    st = oracle_state(mem)
    d = st.dom
    nx, ny, nz = d.nx, d.ny, d.nz
    # node coordinates of the halo layout, index 0 = node -1
    x = d.lower[0] + (np.arange(nx + 3) - 1) * d.dx[0]
    y = d.lower[1] + (np.arange(ny + 3) - 1) * d.dx[1]
    z = d.lower[2] + (np.arange(nz + 3) - 1) * d.dx[2]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    rng = np.random.default_rng(17)
    coef = rng.normal(size=(14, 4))

    def lin(q, P):
        return coef[q, 0] + coef[q, 1] * P[..., 0] + coef[q, 2] * P[..., 1] + coef[q, 3] * P[..., 2]

    G = np.stack([X, Y, Z], axis=-1)
    for l in range(3):
        st.velog[l][:] = lin(l, G)
        st.vtend[l][:] = lin(11 + l, G)
    for l in range(8):
        st.velgradg[l][:] = lin(3 + l, G)
    st.delta_pos[:] = 0.5
    st.grid2par(add=True)
    P = st.position
    scale = np.abs(coef).sum(axis=1).max() * 4
    for l in range(3):
        assert np.abs(st.delta_pos[:, l] - (0.5 + lin(l, P))).max() < 1e-13 * scale
        assert np.abs(st.delta_vor[:, l] - lin(11 + l, P)).max() < 1e-13 * scale
    for l in range(8):
        assert np.abs(st.strain[:, l] - lin(3 + l, P)).max() < 1e-13 * scale
    # without add the deltas restart from zero; strain is reset either way
    st.grid2par()
    assert np.abs(st.delta_pos[:, 0] - lin(0, P)).max() < 1e-13 * scale
