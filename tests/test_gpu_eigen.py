"""The device eigen-solver (the Scherzinger-Dohrmann solve every par2grid,
push, split and merge applies to the parcels' B) on its own, through
epic_diagonalise:

- the reference's golden set unit-tests/3d/{B,D,V}.asc (1000 matrices, numpy
  eigh), with test_scherzinger_2.f90's metric and tolerance (1e-12);
- random triaxial shapes and near-degenerate pairs against the FP64 oracle:
  eigenvalues to 1e-13 of the largest, eigenvectors of isolated eigenvalues
  to 1e-9 (up to sign), and for a
  (near-)degenerate pair the projector onto the pair's plane;
- the diagonal shortcut (off-diagonal sum below 1e-12): V = I, D sorted.
"""
import os

import numpy as np
import pytest

from epic_amd.core import diagonalise

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as fh:
        n = int(fh.readline())
        a = np.loadtxt(fh)
    assert a.shape[0] == n
    return a


def _full(b):
    return np.array([[b[0], b[1], b[2]], [b[1], b[3], b[4]], [b[2], b[4], b[5]]])


def test_device_scherzinger_golden():
    B = _load("B.asc")
    Dg = _load("D.asc")
    Vg = _load("V.asc")
    A = np.stack([_full(b) for b in B])
    D, V = diagonalise(A)
    err = 0.0
    for n in range(A.shape[0]):
        err = max(err, np.abs(Dg[n] - D[n]).sum())
        Q = Vg[n].reshape(3, 3)   # row k = eigenvector k
        for k in range(3):
            q = Q[k].copy()
            if q[0] / V[n, 0, k] < 0:
                q = -q
            err = max(err, np.abs(q - V[n, :, k]).sum())
    print(f"  golden set: max error {err:.2e}", flush=True)
    assert err < 1e-12, err


def _rot(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_device_scherzinger_vs_oracle(orc):
    rng = np.random.default_rng(31)
    mats, kinds = [], []
    for i in range(3000):
        R = _rot(rng)
        if i % 3 == 0:
            d = np.sort(rng.uniform(0.01, 2.0, 3))[::-1]
            kinds.append("triaxial")
        elif i % 3 == 1:
            a = rng.uniform(0.5, 2.0)
            d = np.array([a, a * (1 + 1e-9 * rng.standard_normal()), 0.3])
            kinds.append("pair01")
        else:
            a = rng.uniform(0.5, 2.0)
            d = np.array([2.5, a, a * (1 + 1e-9 * rng.standard_normal())])
            kinds.append("pair12")
        mats.append(R @ np.diag(d) @ R.T)
    A = np.stack(mats)
    A = 0.5 * (A + np.transpose(A, (0, 2, 1)))
    D, V = diagonalise(A)
    for n in range(A.shape[0]):
        d_ref, v_ref = orc.scherzinger_diagonalise(A[n])
        scale = np.abs(d_ref).max()
        assert np.abs(D[n] - d_ref).max() <= 1e-13 * scale, (n, kinds[n], D[n], d_ref)
        iso = {"triaxial": (0, 1, 2), "pair01": (2,), "pair12": (0,)}[kinds[n]]
        for k in iso:   # up to sign, as the golden-set metric
            e = min(np.abs(V[n, :, k] - v_ref[:, k]).max(), np.abs(V[n, :, k] + v_ref[:, k]).max())
            assert e < 1e-9, (n, kinds[n], k)
        if kinds[n] != "triaxial":
            pr = [k for k in range(3) if k not in iso]
            P = sum(np.outer(V[n, :, k], V[n, :, k]) for k in pr)
            Pr = sum(np.outer(v_ref[:, k], v_ref[:, k]) for k in pr)
            assert np.abs(P - Pr).max() < 1e-12, (n, kinds[n])
        # orthonormal and A V = V D
        assert np.abs(V[n].T @ V[n] - np.eye(3)).max() < 1e-13
        assert np.abs(A[n] @ V[n] - V[n] * D[n]).max() < 1e-12 * scale


def test_device_scherzinger_diagonal_shortcut():
    A = np.diag([1.0, 3.0, 2.0])
    A[0, 1] = A[1, 0] = 1e-14
    D, V = diagonalise(A[None])
    assert list(D[0]) == [3.0, 2.0, 1.0]
    assert np.array_equal(np.abs(V[0]), np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], dtype=float))
