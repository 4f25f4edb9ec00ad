"""Seeded synthetic members (parcels + domain) for the parcel-grid core.

Each generator returns a :class:`Member` whose parcel attributes are already
rounded to the precision the device stores them in (FP32 for everything but
the FP64 positions), so that the FP64 reference run and the device run start
from bit-identical inputs.

Workloads (sizes as named):
  * ``beltrami``        -- 32^3 Beltrami flow on [-pi/2, pi/2]^3, 2x2x2 parcel
                           lattice of spheres, vorticity omega = alpha u
                           interpolated from the grid (262,144 parcels);
  * ``rising_thermal``  -- n^3 grid, 8 parcels/cell jittered by +-1/4 cell,
                           triaxial ellipsoids (a/c log-uniform in [1,4],
                           b = sqrt(ac)), Haar-random orientation, k^-3
                           Gaussian random vorticity of rms 0.1, Gaussian
                           buoyancy blob (128^3 -> 16,777,216 parcels);
  * ``moist_member``    -- as rising_thermal plus a humidity plume and a
                           32-blob buoyancy field (256^3 -> 134M parcels);
  * ``ensemble_member`` -- rising-thermal statistics on an uneven grid.
"""
from __future__ import annotations

import dataclasses
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

DEFAULT_SEED = 20261202

# physics defaults (physics.f90:48-71, 170-175)
GRAVITY = 9.80616
L_V = 2.501e6
C_P = 1005.7
THETA_0 = 288.15
Q_0 = 0.015
HEIGHT_C = 1000.0
GLAT = GRAVITY * L_V / (C_P * THETA_0)
LAMBDA_C = 1.0 / HEIGHT_C


@dataclasses.dataclass
class Member:
    nx: int
    ny: int
    nz: int
    lower: tuple
    extent: tuple
    position: np.ndarray      # (n, 3) float64
    vorticity: np.ndarray     # (n, 3)
    B: np.ndarray             # (n, 5)
    volume: np.ndarray        # (n,)
    buoyancy: np.ndarray      # (n,)
    humidity: np.ndarray      # (n,)
    glat: float = GLAT
    lambda_c: float = LAMBDA_C
    q0: float = Q_0
    f_cor: tuple = (0.0, 0.0, 0.0)
    zeta_zero: tuple = (False, False)
    name: str = ""

    @property
    def n(self):
        return self.position.shape[0]

    @property
    def dx(self):
        return tuple(self.extent[i] / (self.nx, self.ny, self.nz)[i] for i in range(3))

    @property
    def vcell(self):
        d = self.dx
        return d[0] * d[1] * d[2]


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def lattice_positions(nx, ny, nz, lower, dx, n_per_dim):
    """Regular lattice in the order of init_regular_positions
    (parcel_init.f90:108-152): cells iz, iy, ix (ix fastest), then inside a
    cell k, j, i (i fastest), offsets (i - 1/2)/n_per_dim."""
    off = (np.arange(1, n_per_dim + 1) - 0.5) / n_per_dim
    iz, iy, ix, k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), off, off, off, indexing="ij")
    pos = np.empty(iz.shape + (3,))
    pos[..., 0] = lower[0] + ix * dx[0] + dx[0] * i
    pos[..., 1] = lower[1] + iy * dx[1] + dx[1] * j
    pos[..., 2] = lower[2] + iz * dx[2] + dx[2] * k
    return pos.reshape(-1, 3)


def _threads():
    return max(1, int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1))


def _chunked(fn, n, min_chunk=1 << 18):
    """Runs fn(lo, hi) over [0, n) in chunks on a thread pool (numpy releases
    the GIL inside its loops); the result does not depend on the chunking."""
    nt = min(_threads(), max(1, n // min_chunk))
    if nt <= 1:
        fn(0, n)
        return
    bounds = np.linspace(0, n, 4 * nt + 1).astype(np.int64)
    with ThreadPoolExecutor(nt) as ex:
        list(ex.map(lambda i: fn(int(bounds[i]), int(bounds[i + 1])), range(len(bounds) - 1)))


def trilinear_at(field, pos, lower, dx, nx, ny):
    """Trilinear interpolation (interpl.f90:17-56) of a node field given on
    nodes (nx, ny, nz+1) [x-periodic, y-periodic] at positions pos (n,3);
    parcels must lie inside [lower_z, upper_z).  field may also be a list of
    such fields; then a list of results is returned."""
    fields = field if isinstance(field, (list, tuple)) else [field]
    # one periodic column and row appended: corner indices need no modulo
    padded = [np.pad(f, ((0, 1), (0, 1), (0, 0)), mode="wrap").reshape(-1) for f in fields]
    sy, sx = fields[0].shape[2], (ny + 1) * fields[0].shape[2]
    outs = [np.empty(pos.shape[0]) for _ in fields]
    low, dxa = np.asarray(lower), np.asarray(dx)

    def run(lo, hi):
        xyz = (pos[lo:hi] - low) / dxa
        ijk = np.floor(xyz).astype(np.int64)
        p = xyz - ijk
        ijk[:, 0] %= nx
        ijk[:, 1] %= ny
        base = ijk[:, 0] * sx + ijk[:, 1] * sy + ijk[:, 2]
        acc = [np.zeros(hi - lo) for _ in fields]
        for kk in (0, 1):
            wz = p[:, 2] if kk else 1 - p[:, 2]
            for jj in (0, 1):
                wy = p[:, 1] if jj else 1 - p[:, 1]
                for ii in (0, 1):
                    wx = p[:, 0] if ii else 1 - p[:, 0]
                    w = wz * wy * wx
                    idx = base + (ii * sx + jj * sy + kk)
                    for a, f in zip(acc, padded):
                        a += w * f[idx]
        for o, a in zip(outs, acc):
            o[lo:hi] = a

    _chunked(run, pos.shape[0])
    return outs if isinstance(field, (list, tuple)) else outs[0]


def _rotation(q):
    """Rotation matrices (m, 3, 3) of unit quaternions q (m, 4)."""
    w, x, y, z = q.T
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def haar_rotations(rng, n):
    """n Haar-random rotation matrices from normalised Gaussian quaternions."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return _rotation(q)


def triaxial_B(rng, volume, lam_lo=1.0, lam_hi=4.0):
    """B = R diag(a^2, b^2, c^2) R^T with a/c log-uniform in [lam_lo, lam_hi],
    b = sqrt(ac) and abc = 3V/(4 pi); returns (n,5) = B11 B12 B13 B22 B23."""
    n = volume.shape[0]
    lam = np.exp(rng.uniform(math.log(lam_lo), math.log(lam_hi), n))
    q = rng.standard_normal((n, 4))
    out = np.empty((n, 5))

    def run(lo, hi):
        qq = q[lo:hi] / np.linalg.norm(q[lo:hi], axis=1, keepdims=True)
        R = _rotation(qq)
        r = (3.0 * volume[lo:hi] / (4.0 * math.pi)) ** (1.0 / 3.0)
        a = r * np.sqrt(lam[lo:hi])
        c = r / np.sqrt(lam[lo:hi])
        b = r
        d2 = np.stack([a * a, b * b, c * c], axis=1)
        Bf = np.einsum("nij,nj,nkj->nik", R, d2, R)
        out[lo:hi] = np.stack([Bf[:, 0, 0], Bf[:, 0, 1], Bf[:, 0, 2], Bf[:, 1, 1], Bf[:, 1, 2]], axis=1)

    _chunked(run, n)
    return out


def gaussian_random_field(rng, nx, ny, nz, slope=-3.0, rms=0.1):
    """Periodic Gaussian random field with shell spectrum E(k) ~ k^slope on an
    (nx, ny, 2nz) box; returns node values (nx, ny, nz+1) scaled to rms."""
    nz2 = 2 * nz
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    kz = np.fft.rfftfreq(nz2) * nz2
    K = np.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.zeros_like(K)
    nzr = K > 0
    amp[nzr] = K[nzr] ** ((slope - 2.0) / 2.0)   # per-mode variance E(k)/k^2
    noise = rng.standard_normal(K.shape) + 1j * rng.standard_normal(K.shape)
    f = np.fft.irfftn(noise * amp, s=(nx, ny, nz2), axes=(0, 1, 2))[:, :, : nz + 1]
    f -= f.mean()
    return f * (rms / np.sqrt(np.mean(f * f)))


def _lattice_member(nx, ny, nz, lower, extent, n_per_dim, rng, jitter):
    dx = tuple(extent[i] / (nx, ny, nz)[i] for i in range(3))
    pos = lattice_positions(nx, ny, nz, lower, dx, n_per_dim)
    if jitter:
        pos = pos + rng.uniform(-jitter, jitter, pos.shape) * np.asarray(dx)
    return pos, dx


def beltrami(n=32, k=1, l=1, m=1, dtype_round=True):
    """Beltrami member: spheres on a 2x2x2 lattice, omega = alpha u sampled
    on the grid (beltrami_3d.f90:80-122) and interpolated to the parcels
    (parcel_init.f90:170-203)."""
    lower = (-0.5 * math.pi,) * 3
    extent = (math.pi,) * 3
    pos, dx = _lattice_member(n, n, n, lower, extent, 2, None, 0.0)
    vcell = dx[0] * dx[1] * dx[2]
    vol = np.full(pos.shape[0], vcell / 8.0)
    abc = 0.75 * vol / math.pi
    l23 = abc ** (2.0 / 3.0)
    B = np.zeros((pos.shape[0], 5))
    B[:, 0] = l23
    B[:, 3] = l23
    alpha = math.sqrt(k * k + l * l + m * m)
    fk2l2 = alpha / (k * k + l * l)
    x = lower[0] + np.arange(n) * dx[0]
    y = lower[1] + np.arange(n) * dx[1]
    z = lower[2] + np.arange(n + 1) * dx[2]
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    cosmz, sinmz = np.cos(m * Z), np.sin(m * Z)
    s, c = np.sin(k * X + l * Y), np.cos(k * X + l * Y)
    om = [fk2l2 * (k * m * sinmz - l * alpha * cosmz) * s,
          fk2l2 * (l * m * sinmz + k * alpha * cosmz) * s,
          alpha * cosmz * c]
    vor = np.stack([trilinear_at(f, pos, lower, dx, n, n) for f in om], axis=1)
    # set_zeta_boundary_flag (parameters.f90:168-217) on the gridded zeta
    zeta = om[2]
    rms_int = math.sqrt(np.sum(zeta[:, :, 1:n] ** 2) / (n * n) / (n - 1))
    thres = 1.0e-3 * rms_int + np.finfo(float).eps
    zz = (math.sqrt(np.sum(zeta[:, :, 0] ** 2) / (n * n)) < thres,
          math.sqrt(np.sum(zeta[:, :, n] ** 2) / (n * n)) < thres)
    r = _f32 if dtype_round else (lambda a: a)
    N = pos.shape[0]
    return Member(n, n, n, lower, extent, pos, r(vor), r(B), r(vol), np.zeros(N), np.zeros(N),
                  zeta_zero=zz, name=f"beltrami{n}")


def rising_thermal(n=128, seed=DEFAULT_SEED, ny=None, nz=None, humidity_plume=False, n_blobs=0,
                   dtype_round=True, log=None):
    """Rising-thermal member on an n x ny x nz grid over [-pi/2, pi/2]^3-like
    box (cell-isotropic extent pi * (n, ny, nz)/n), 8 parcels per cell.
    log(str), if given, is called after each stage of the generation."""
    say = log or (lambda s: None)
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    rng = np.random.default_rng(seed)
    ext0 = math.pi
    extent = (ext0, ext0 * ny / n, ext0 * nz / n)
    lower = tuple(-0.5 * e for e in extent)
    pos, dx = _lattice_member(n, ny, nz, lower, extent, 2, rng, 0.25)
    N = pos.shape[0]
    say(f"lattice: {N} parcels")
    vcell = dx[0] * dx[1] * dx[2]
    vol = np.full(N, vcell / 8.0)
    B = triaxial_B(rng, vol)
    say("shape matrices")
    grf = [gaussian_random_field(rng, n, ny, nz) for _ in range(3)]
    vor = np.stack(trilinear_at(grf, pos, lower, dx, n, ny), axis=1)
    say("vorticity")
    L = np.asarray(extent)
    rel = (pos - np.asarray(lower)) / L
    if n_blobs:
        buoy = np.zeros(N)
        amps = rng.normal(0.0, 0.5, n_blobs)
        centres = rng.uniform(0.0, 1.0, (n_blobs, 3))
        for b, (a, c0) in enumerate(zip(amps, centres)):
            r2 = np.sum(((rel - c0) * L) ** 2, axis=1)
            buoy += a * np.exp(-r2 / (0.1 * L[0]) ** 2)
            if b % 8 == 7:
                say(f"buoyancy blob {b + 1} of {n_blobs}")
    else:
        r2 = np.sum(((rel - np.array([0.5, 0.5, 0.25])) * L) ** 2, axis=1)
        buoy = np.exp(-r2 / (0.2 * L[0]) ** 2)
    hum = np.zeros(N)
    if humidity_plume:
        r2 = np.sum(((rel - np.array([0.5, 0.5, 0.25])) * L) ** 2, axis=1)
        R = 0.15 * L[0]
        inside = r2 < R * R
        hum[inside] = rng.uniform(0.0, 1.0, inside.sum()) * np.exp(-r2[inside] / (R * R))
    r = _f32 if dtype_round else (lambda a: a)
    return Member(n, ny, nz, lower, extent, pos, r(vor), r(B), r(vol), r(buoy), r(hum),
                  name=f"rising_thermal{n}x{ny}x{nz}")


def moist_member(n=256, seed=DEFAULT_SEED, log=None):
    return rising_thermal(n, seed=seed, humidity_plume=True, n_blobs=32, log=log)


def ensemble_member(m, seed0=DEFAULT_SEED, sizes_xy=(32, 48, 64, 96), sizes_z=(32, 64)):
    rng = np.random.default_rng(seed0 + 7919 * m)
    nx = int(rng.choice(sizes_xy))
    ny = int(rng.choice(sizes_xy))
    nz = int(rng.choice(sizes_z))
    mem = rising_thermal(nx, seed=seed0 + m, ny=ny, nz=nz)
    mem.name = f"member{m}_{nx}x{ny}x{nz}"
    return mem
