"""Oracle transforms and inversion against the reference's own known-answer
tests (unit-tests/3d/test_mpi_*.f90) and textbook sums."""
import numpy as np
import pytest

PI = np.pi


def _grid(orc, dom, fn, zlo=0, zhi=None):
    """Evaluate fn(x, y, z) on interior nodes into a halo array [ix+1, iy+1, iz+1]."""
    nx, ny, nz = dom.nx, dom.ny, dom.nz
    zhi = nz if zhi is None else zhi
    x = dom.lower[0] + np.arange(nx) * dom.dx[0]
    y = dom.lower[1] + np.arange(ny) * dom.dx[1]
    z = dom.lower[2] + np.arange(zlo, zhi + 1) * dom.dx[2]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    a = np.zeros((nx + 3, ny + 3, nz + 3))
    a[1:nx + 1, 1:ny + 1, zlo + 1:zhi + 2] = fn(X, Y, Z)
    return a


def test_forfft_convention_and_roundtrip(orc):
    rng = np.random.default_rng(0)
    for n in (8, 12, 16, 24, 32, 48, 96, 128, 256, 30):
        x = rng.standard_normal(n)
        h = orc.forfft(x)
        X = np.fft.fft(x) / np.sqrt(n)
        assert np.allclose(h[: n // 2 + 1], X[: n // 2 + 1].real, atol=1e-13)
        k = np.arange(1, (n - 1) // 2 + 1)
        assert np.allclose(h[n - k], X[k].imag, atol=1e-13)
        assert np.abs(orc.revfft(h) - x).max() < 1e-13


@pytest.mark.parametrize("n", [16, 32, 48, 64, 96, 128, 256])
def test_dst_dct_textbook_and_involution(orc, n):
    """DST-I / DCT-I with the sqrt(2/n) normalisation, self-inverse
    (stafft.f90:410-555; test_mpi_dst.f90:86 tolerance 1e-14)."""
    rng = np.random.default_rng(n)
    j = np.arange(1, n)
    S = np.sqrt(2.0 / n) * np.sin(PI * np.outer(j, j) / n)
    x = rng.standard_normal(n)
    y = orc.dst(x)
    assert y[-1] == 0.0
    assert np.abs(y[:-1] - S @ x[:-1]).max() < 1e-13
    assert np.abs(orc.dst(y)[:-1] - x[:-1]).max() < 1e-13
    k = np.arange(n + 1)
    w = np.ones(n + 1)
    w[0] = w[-1] = 0.5
    Cm = np.sqrt(2.0 / n) * np.cos(PI * np.outer(k, k) / n) * w[None, :]
    x = rng.standard_normal(n + 1)
    y = orc.dct(x)
    assert np.abs(y - Cm @ x).max() < 1e-13
    assert np.abs(orc.dct(y) - x).max() < 1e-13


def test_fft0_roundtrip(orc):
    """test_mpi_fft0.f90: fp = cos(4x) + sin(y) + cos(z), p2s then s2p, < 1e-14."""
    d = orc.make_dom(32, 64, 128, (-PI, 0.5 * PI, 0.0), (2 * PI, 4 * PI, 8 * PI))
    st = orc.State(d, 0)
    fp = _grid(orc, d, lambda x, y, z: np.cos(4 * x) + np.sin(y) + np.cos(z))
    fs = st.fftxyp2s(fp)
    back = st.fftxys2p(fs)
    assert np.abs(back[1:-2, 1:-2, 1:-1] - fp[1:-2, 1:-2, 1:-1]).max() < 1e-14


def test_dst_fftsine_roundtrip(orc):
    """test_mpi_dst.f90: cos(4x) sin(y) sin(z); p2s, fftsine twice, s2p, < 1e-14."""
    d = orc.make_dom(32, 64, 128, (-PI, 0.5 * PI, 0.0), (2 * PI, 4 * PI, 8 * PI))
    st = orc.State(d, 0)
    fp = _grid(orc, d, lambda x, y, z: np.cos(4 * x) * np.sin(y) * np.sin(z))
    fs = st.fftxyp2s(fp)
    nz = d.nz
    for _ in range(2):
        for ix in range(d.nx):
            for iy in range(d.ny):
                fs[ix, iy, 1:nz + 1] = orc.dst(fs[ix, iy, 1:nz + 1])
    back = st.fftxys2p(fs)
    assert np.abs(back[1:-2, 1:-2, 1:-1] - fp[1:-2, 1:-2, 1:-1]).max() < 1e-14


def test_diffx_diffy(orc):
    """test_mpi_diffx.f90 / test_mpi_diffy.f90: d/dx cos(4x) = -4 sin(4x) (1e-12)."""
    d = orc.make_dom(64, 128, 64, (-PI, -PI, -PI), (2 * PI, 2 * PI, 2 * PI))
    st = orc.State(d, 0)
    fp = _grid(orc, d, lambda x, y, z: np.cos(4 * x) + 0 * y)
    out = st.fftxys2p(st.diffx(st.fftxyp2s(fp)))
    want = _grid(orc, d, lambda x, y, z: -4 * np.sin(4 * x) + 0 * y)
    assert np.abs(out - want)[1:-2, 1:-2, 1:-1].max() < 1e-12
    fp = _grid(orc, d, lambda x, y, z: np.cos(4 * y) + 0 * x)
    out = st.fftxys2p(st.diffy(st.fftxyp2s(fp)))
    want = _grid(orc, d, lambda x, y, z: -4 * np.sin(4 * y) + 0 * x)
    assert np.abs(out - want)[1:-2, 1:-2, 1:-1].max() < 1e-12


def _taylor_green(A, B, Cc, a, b, c):
    vel = [lambda x, y, z: A * np.cos(a * x) * np.sin(b * y) * np.sin(c * z),
           lambda x, y, z: B * np.sin(a * x) * np.cos(b * y) * np.sin(c * z),
           lambda x, y, z: Cc * np.sin(a * x) * np.sin(b * y) * np.cos(c * z)]
    vor = [lambda x, y, z: (b * Cc - c * B) * np.sin(a * x) * np.cos(b * y) * np.cos(c * z),
           lambda x, y, z: (c * A - a * Cc) * np.cos(a * x) * np.sin(b * y) * np.cos(c * z),
           lambda x, y, z: (a * B - b * A) * np.cos(a * x) * np.cos(b * y) * np.sin(c * z)]
    return vel, vor


def test_vor2vel_taylor_green(orc):
    """test_mpi_vor2vel.f90: 64^3, A=B=1, C=-2, a=b=c=1, error < 1e-14 (:102)."""
    d = orc.make_dom(64, 64, 64, (-PI, -PI, -0.5 * PI), (2 * PI, 2 * PI, 2 * PI))
    st = orc.State(d, 0)
    vel, vor = _taylor_green(1.0, 1.0, -2.0, 1.0, 1.0, 1.0)
    for c in range(3):
        st.vortg[c] = _grid(orc, d, vor[c], zlo=-1, zhi=d.nz + 1)
    st.vor2vel()
    err = 0.0
    for c in range(3):
        want = _grid(orc, d, vel[c])
        err = max(err, np.abs(st.velog[c] - want)[1:-2, 1:-2, 1:-1].max())
    assert err < 1e-14, err


def test_vel2vgrad_taylor_green(orc):
    """test_mpi_vel2vgrad.f90: 32x64x128, A=1,a=2,B=-1,b=1,C=1,c=-1, < 2e-14 (:136)."""
    d = orc.make_dom(32, 64, 128, (-PI, -2 * PI, -4 * PI), (2 * PI, 4 * PI, 8 * PI))
    st = orc.State(d, 0)
    A, a, B, b, Cc, c = 1.0, 2.0, -1.0, 1.0, 1.0, -1.0
    vel, vor = _taylor_green(A, B, Cc, a, b, c)
    for k in range(3):
        st.vortg[k] = _grid(orc, d, vor[k])
    svel = np.stack([st.fftxyp2s(_grid(orc, d, vel[k])) for k in range(3)])
    st.vel2vgrad(svel)
    s = lambda f: f  # noqa: E731
    dudx = lambda x, y, z: -a * A * np.sin(a * x) * np.sin(b * y) * np.sin(c * z)
    dudy = lambda x, y, z: b * A * np.cos(a * x) * np.cos(b * y) * np.sin(c * z)
    dvdy = lambda x, y, z: -b * B * np.sin(a * x) * np.sin(b * y) * np.sin(c * z)
    dwdx = lambda x, y, z: a * Cc * np.cos(a * x) * np.sin(b * y) * np.cos(c * z)
    dwdy = lambda x, y, z: b * Cc * np.sin(a * x) * np.cos(b * y) * np.cos(c * z)
    want = [dudx, dudy, lambda x, y, z: dwdx(x, y, z) + vor[1](x, y, z),
            lambda x, y, z: dudy(x, y, z) + vor[2](x, y, z), dvdy,
            lambda x, y, z: dwdy(x, y, z) - vor[0](x, y, z), dwdx, dwdy]
    err = 0.0
    for k in range(8):
        err = max(err, np.abs(st.velgradg[k] - _grid(orc, d, s(want[k])))[1:-2, 1:-2, 1:-1].max())
    assert err < 2e-14, err


@pytest.mark.parametrize("k,l,m", [(1, 1, 1), (2, 3, 3), (3, 1, 1)])
def test_vtend_beltrami(orc, k, l, m):
    """test_mpi_vtend.f90: Beltrami flow on 128^3, error < 1.2 (:124); a subset of
    the reference's (k,l,m) grid to bound CPU time."""
    d = orc.make_dom(128, 128, 128, (-0.5 * PI,) * 3, (PI,) * 3)
    st = orc.State(d, 0)
    alpha = np.sqrt(k * k + l * l + m * m)
    fk2l2 = 1.0 / (k * k + l * l)
    u = lambda x, y, z: fk2l2 * (k * m * np.sin(m * z) - l * alpha * np.cos(m * z)) * np.sin(k * x + l * y)
    v = lambda x, y, z: fk2l2 * (l * m * np.sin(m * z) + k * alpha * np.cos(m * z)) * np.sin(k * x + l * y)
    w = lambda x, y, z: np.cos(m * z) * np.cos(k * x + l * y)
    for c, fn in enumerate((u, v, w)):
        st.velog[c] = _grid(orc, d, fn, zlo=-1, zhi=d.nz + 1)
        st.vortg[c] = alpha * st.velog[c]
    st.vorticity_tendency()
    ref = [lambda x, y, z: alpha * k * m * m * fk2l2 * np.sin(k * x + l * y) * np.cos(k * x + l * y),
           lambda x, y, z: alpha * l * m * m * fk2l2 * np.sin(k * x + l * y) * np.cos(k * x + l * y),
           lambda x, y, z: -alpha * m * np.sin(m * z) * np.cos(m * z)]
    err = max(np.abs(st.vtend[c] - _grid(orc, d, ref[c]))[1:-2, 1:-2, 1:-1].max() for c in range(3))
    assert err < 1.2, err
    # interior (away from the one-sided wall differences) is second-order accurate
    inner = max(np.abs(st.vtend[c] - _grid(orc, d, ref[c]))[1:-2, 1:-2, 3:-3].max() for c in range(3))
    assert inner < 0.05 * alpha * m * m, inner
