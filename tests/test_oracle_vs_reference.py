"""The oracle's transforms and eigen-solver against the reference's own
Fortran (stafft.f90, scherzinger.f90, compiled from the reference sources by
oracle/ref/Makefile into oracle/_ref/libref.so).  Skipped where that build is
absent (it needs the reference sources and amdflang)."""
import ctypes as C
import os

import numpy as np
import pytest

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="reference leaf modules not built")


@pytest.fixture(scope="module")
def ref():
    L = C.CDLL(REF)
    P = C.POINTER(C.c_double)
    for f in ("ref_forfft", "ref_revfft", "ref_dct", "ref_dst"):
        getattr(L, f).argtypes = [C.c_int, P]
        getattr(L, f).restype = None
    L.ref_scherzinger.argtypes = [P, P, P]
    L.ref_scherzinger.restype = None
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.mark.parametrize("n", [8, 12, 16, 24, 30, 32, 48, 64, 96, 128, 256])
def test_transforms_match_reference(orc, ref, n):
    rng = np.random.default_rng(n)
    for name, size in (("forfft", n), ("revfft", n), ("dct", n + 1), ("dst", n)):
        x = rng.standard_normal(size)
        want = x.copy()
        if name in ("dct",):
            getattr(ref, "ref_" + name)(n, _p(want))
        else:
            getattr(ref, "ref_" + name)(n, _p(want))
        got = getattr(orc, name)(x)
        scale = max(1.0, np.abs(want).max())
        if name == "dst":   # the reference leaves x(n) = 0
            assert got[-1] == 0.0 and want[-1] == 0.0
        assert np.abs(got - want).max() < 1e-14 * scale, (name, np.abs(got - want).max())


def _ref_eig(ref, A):
    a = np.asfortranarray(A.astype(np.float64))
    D = np.zeros(3)
    V = np.zeros((3, 3), order="F")
    ref.ref_scherzinger(a.ctypes.data_as(C.POINTER(C.c_double)), _p(D), V.ctypes.data_as(C.POINTER(C.c_double)))
    return D, np.array(V)


def test_scherzinger_matches_reference(orc, ref):
    """Random, golden-set-like, diagonal and near-degenerate (spheroidal) shapes:
    same eigenvalues, same eigenvector basis and signs as the reference."""
    rng = np.random.default_rng(17)
    mats = []
    for _ in range(2000):
        b = rng.uniform(-100, 100, (3, 3))
        mats.append((b + b.T) / 2)
    for _ in range(2000):   # ellipsoid shapes incl. near-spheres and spheroids
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        kind = rng.integers(4)
        a = 1.0
        if kind == 0:
            d = [a, rng.uniform(0.2, 1), rng.uniform(0.05, 0.2)]
        elif kind == 1:
            d = [a, a * (1 + 1e-9 * rng.standard_normal()), 0.3]        # prolate-ish degenerate pair
        elif kind == 2:
            d = [a, 0.5, 0.5 * (1 + 1e-9 * rng.standard_normal())]
        else:
            d = [a, a, a]
        mats.append(R @ np.diag(np.square(d)) @ R.T)
    mats.append(np.diag([3.0, 1.0, 2.0]))
    # Isolated eigenvalues: same eigenvector with the same sign.  Inside a
    # (near-)degenerate pair the basis is fixed by rounding only (also between
    # compilers of the reference), so the pair is compared by its projector.
    worst_d = worst_v = worst_p = 0.0
    for A in mats:
        Dr, Vr = _ref_eig(ref, A)
        Do, Vo = orc.scherzinger_diagonalise(A)
        sc = max(1.0, np.abs(Dr).max())
        worst_d = max(worst_d, np.abs(Do - Dr).max() / sc)
        gap = [min(abs(Dr[i] - Dr[j]) for j in range(3) if j != i) / sc for i in range(3)]
        for i in range(3):
            if gap[i] > 1e-6:
                worst_v = max(worst_v, np.abs(Vo[:, i] - Vr[:, i]).max())
        for i in range(3):
            for j in range(i + 1, 3):
                if abs(Dr[i] - Dr[j]) / sc <= 1e-6 and all(gap[k] > 1e-6 for k in range(3) if k not in (i, j)):
                    Pr = np.outer(Vr[:, i], Vr[:, i]) + np.outer(Vr[:, j], Vr[:, j])
                    Po = np.outer(Vo[:, i], Vo[:, i]) + np.outer(Vo[:, j], Vo[:, j])
                    worst_p = max(worst_p, np.abs(Po - Pr).max())
    assert worst_d < 1e-13, worst_d
    assert worst_v < 1e-9, worst_v
    assert worst_p < 1e-12, worst_p
