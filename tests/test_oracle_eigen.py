"""Oracle eigen-solver and ellipsoid geometry against the reference's golden
vectors (unit-tests/3d/{B,D,V}.asc, test_scherzinger_2.f90:47 tolerance 1e-12)
and against closed forms."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as fh:
        n = int(fh.readline())
        a = np.loadtxt(fh)
    assert a.shape[0] == n
    return a


def test_scherzinger_golden(orc):
    B = _load("B.asc")
    D = _load("D.asc")
    V = _load("V.asc")
    err = 0.0
    for n in range(B.shape[0]):
        b = B[n]
        A = np.array([[b[0], b[1], b[2]], [b[1], b[3], b[4]], [b[2], b[4], b[5]]])
        d, v = orc.scherzinger_diagonalise(A)
        err = max(err, np.abs(D[n] - d).sum())
        Q = V[n].reshape(3, 3)   # row j = eigenvector j
        for k in range(3):
            q = Q[k].copy()
            if q[0] / v[0, k] < 0:
                q = -q
            err = max(err, np.abs(q - v[:, k]).sum())
    assert err < 1e-12, err


def test_scherzinger_diagonal_shortcut(orc):
    # off-diagonal sum below atol: V = I sorted, D = diagonal sorted (scherzinger.f90:41-56)
    A = np.diag([1.0, 3.0, 2.0])
    A[0, 1] = A[1, 0] = 1e-14
    d, v = orc.scherzinger_diagonalise(A)
    assert list(d) == [3.0, 2.0, 1.0]
    assert np.array_equal(np.abs(v), np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], dtype=float))


def _rot(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_B33_volume_constraint(orc):
    """det B = (3V/4pi)^2 with B33 re-derived (parcel_ellipsoid.f90:383-402)."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        a, b, c = np.sort(rng.uniform(0.1, 1.0, 3))[::-1]
        R = _rot(rng)
        Bf = R @ np.diag([a * a, b * b, c * c]) @ R.T
        V = 4.0 / 3.0 * np.pi * a * b * c
        B5 = np.array([Bf[0, 0], Bf[0, 1], Bf[0, 2], Bf[1, 1], Bf[1, 2]])
        b33 = orc.get_B33(B5, V)
        assert abs(b33 - Bf[2, 2]) < 1e-12 * max(1.0, abs(Bf[2, 2]))


def test_B33_degenerate_raises(orc):
    with pytest.raises(orc.OracleError) as e:
        orc.get_B33(np.array([1.0, 1.0, 0.0, 1.0, 0.0]), 1.0)
    assert e.value.code == -2


def test_support_points_geometry(orc):
    """Support points lie on the ellipsoid at radius rho=sqrt(2/5) in the
    (a, b) plane: x^T B^-1 x = rho^2 for x = point - centre, and their mean is
    the centre (parcel_ellipsoid.f90:29-34, 216-260)."""
    rng = np.random.default_rng(3)
    for _ in range(100):
        a, b, c = 1.0, rng.uniform(0.3, 0.9), rng.uniform(0.05, 0.25)
        R = _rot(rng)
        Bf = R @ np.diag([a * a, b * b, c * c]) @ R.T
        V = 4.0 / 3.0 * np.pi * a * b * c
        B5 = np.array([Bf[0, 0], Bf[0, 1], Bf[0, 2], Bf[1, 1], Bf[1, 2]])
        pos = rng.standard_normal(3)
        ve, vt, pts = orc.get_points(pos, B5, V)
        assert np.allclose(pts.mean(axis=0), pos, atol=1e-13)
        # the two semi-axis vectors: |Veta| = rho sqrt(a^2 - c^2), |Vtau| = rho sqrt(b^2 - c^2)
        rho = np.sqrt(0.4)
        assert abs(np.linalg.norm(ve) - rho * np.sqrt(a * a - c * c)) < 1e-12
        assert abs(np.linalg.norm(vt) - rho * np.sqrt(b * b - c * c)) < 1e-12
        assert abs(np.dot(ve, vt)) < 1e-12
        # Veta is along the major axis R[:, 0] (up to sign)
        assert abs(abs(np.dot(ve / np.linalg.norm(ve), R[:, 0])) - 1.0) < 1e-10


def test_dBdt_is_BLt_plus_LB(orc):
    """dB/dt = L B + B L^T for the velocity gradient L with dw/dz = -(du/dx+dv/dy)
    (parcel_ellipsoid.f90:466-507; the 5 stored components)."""
    rng = np.random.default_rng(11)
    for _ in range(50):
        a, b, c = np.sort(rng.uniform(0.2, 1.0, 3))[::-1]
        R = _rot(rng)
        Bf = R @ np.diag([a * a, b * b, c * c]) @ R.T
        V = 4.0 / 3.0 * np.pi * a * b * c
        B5 = np.array([Bf[0, 0], Bf[0, 1], Bf[0, 2], Bf[1, 1], Bf[1, 2]])
        s = rng.standard_normal(8)   # dudx dudy dudz dvdx dvdy dvdz dwdx dwdy
        L = np.array([[s[0], s[1], s[2]], [s[3], s[4], s[5]], [s[6], s[7], -(s[0] + s[4])]])
        full = L @ Bf + Bf @ L.T
        want = np.array([full[0, 0], full[0, 1], full[0, 2], full[1, 1], full[1, 2]])
        got = orc.get_dBdt(B5, s, V)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-13)
