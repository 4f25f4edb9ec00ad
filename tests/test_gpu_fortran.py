"""The Fortran ISO_C_BINDING module drives the same GPU core as the C-ABI.

The example driver (epic_amd/fortran/drive_beltrami.f90) builds a 16^3
Beltrami member with analytic vorticity at the parcel centres and runs two
iterations of EPIC's run loop (apply_vortcor, a fixed-dt RK4 ``ls_rk_step``,
parcel_merge, parcel_split, two rounds of apply_laplace / apply_gradient),
then prints checksums of the final parcel state and of par2grid's volume
field.  The same member run through the
C-ABI from Python must give the same parcel state (the device computation is
deterministic and identical), so the checksums agree to summation-order
rounding; a transposed or misread argument in the Fortran interfaces would not.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from epic_amd import build as B
from epic_amd import workloads as W
from epic_amd.core import Member

pytestmark = pytest.mark.gpu


def _member_inputs(n=16):
    lower = (-0.5 * math.pi,) * 3
    extent = (math.pi,) * 3
    dx = tuple(e / n for e in extent)
    pos = W.lattice_positions(n, n, n, lower, dx, 2)
    vol = np.full(pos.shape[0], dx[0] * dx[1] * dx[2] / 8.0)
    l23 = (0.75 * vol / math.pi) ** (2.0 / 3.0)
    Bm = np.zeros((pos.shape[0], 5))
    Bm[:, 0] = l23
    Bm[:, 3] = l23
    alpha = math.sqrt(3.0)
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    vor = np.stack([alpha / 2 * (np.sin(z) - alpha * np.cos(z)) * np.sin(x + y),
                    alpha / 2 * (np.sin(z) + alpha * np.cos(z)) * np.sin(x + y),
                    alpha * np.cos(z) * np.cos(x + y)], axis=1)
    return lower, extent, pos, vor, Bm, vol


@pytest.mark.skipif(not os.path.exists(B.FORTRAN_DRIVER), reason="Fortran driver not built (no amdflang)")
def test_fortran_driver_matches_c_abi():
    print("fortran driver: beltrami 16^3, two run-loop iterations", flush=True)
    res = subprocess.run([B.FORTRAN_DRIVER], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    line = [l for l in res.stdout.splitlines() if l.startswith("checksums")][0]
    got = np.array([float(v) for v in line.split()[1:]])

    n = 16
    lower, extent, pos, vor, Bm, vol = _member_inputs(n)
    m = Member(n, n, n, lower, extent, zeta_zero=(True, True), max_parcels=2 * pos.shape[0])
    m.set_parcels(pos, vor, Bm, vol, np.zeros(len(vol)), np.zeros(len(vol)))
    m.init_parcel_correction()
    t = 0.0
    for _ in range(2):
        m.apply_vortcor()
        t = m.ls_rk_step(t, 0.01, 0.0)
        assert m.parcel_merge() == 0
        assert m.parcel_split() == 0
        for it in (1, 2):
            m.apply_laplace(l_reuse=it > 1)
            m.apply_gradient(1.8, 0.5, l_reuse=True)
    p = m.get_parcels()
    m.par2grid()
    volg = m.get_field("volg")
    # volg is indexed [ix + 1, iy + 1, iz + 1]: nodes x 0..nx-1, y 0..ny-1, z 0..nz
    want = np.array([np.abs(p["position"]).sum(), np.abs(p["vorticity"]).sum(), p["B"][:, 0].sum(),
                     p["B"][:, 3].sum(), p["volume"].sum(), volg[1:n + 1, 1:n + 1, 1:n + 2].sum()])
    m.close()
    assert abs(t - 0.02) < 1e-15
    np.testing.assert_allclose(got, want, rtol=1e-12)
